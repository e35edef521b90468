"""Host-side handle on libhelio_hip.so: a context, device buffers and the
pass-level calls, mirroring the reference passes.

``LightCullPass`` and ``DeferredLightPass`` keep the names, the published
frame slots and the prepare/execute split of
crates/passes/3d/helio-pass-light-cull/src/lib.rs:41-363 and
crates/passes/3d/helio-pass-deferred-light/src/lib.rs:54-1230; the work of
``execute`` is one C-ABI call each.  Missing native code raises instead of
falling back: there is no CPU path here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi
from .abi import HelioError

_LIB: abi.Lib | None = None


def lib() -> abi.Lib:
    global _LIB
    if _LIB is None:
        _LIB = abi.Lib()
    return _LIB


class DeviceArray:
    """A device allocation made through hl_alloc."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx.lib.call("hl_alloc", ctx.h, max(self.nbytes, 16), C.byref(p))
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes, (a.nbytes, self.nbytes)
        if a.nbytes:
            self.ctx.lib.call("hl_memcpy_htod", self.ctx.h, self.ptr, a.ctypes.data, a.nbytes)
        return self

    def download(self, dtype, shape=None, nbytes=None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes if nbytes is None else nbytes
        out = np.empty(n // dt.itemsize, dtype=dt)
        if out.nbytes:
            self.ctx.lib.call("hl_memcpy_dtoh", self.ctx.h, out.ctypes.data, self.ptr, out.nbytes)
        return out.reshape(shape) if shape is not None else out

    def fill_u32(self, value: int, count: int | None = None) -> None:
        n = self.nbytes // 4 if count is None else count
        self.ctx.lib.call("hl_memset_d32", self.ctx.h, self.ptr, value & 0xFFFFFFFF, n)

    def free(self) -> None:
        if self.ptr:
            self.ctx.lib.call("hl_free", self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.ctx.lib.raw["hl_free"](self.ctx.h, self.ptr)
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0):
        self.lib = lib()
        h = C.c_void_p()
        self.lib.call("hl_context_create", device, C.byref(h))
        self.h = h.value

    def close(self) -> None:
        if self.h:
            self.lib.call("hl_context_destroy", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def alloc(self, nbytes: int) -> DeviceArray:
        return DeviceArray(self, nbytes)

    def upload(self, arr: np.ndarray) -> DeviceArray:
        a = np.ascontiguousarray(arr)
        return DeviceArray(self, a.nbytes).upload(a)

    def synchronize(self) -> None:
        self.lib.call("hl_synchronize", self.h)

    def last_error(self) -> str:
        return self.lib.raw["hl_last_error"](self.h).decode()

    # events
    def event(self):
        e = C.c_void_p()
        self.lib.call("hl_event_create", self.h, C.byref(e))
        return e.value

    def record(self, ev) -> None:
        self.lib.call("hl_event_record", self.h, ev)

    def elapsed_ms(self, a, b) -> float:
        ms = C.c_float()
        self.lib.call("hl_event_elapsed_ms", a, b, C.byref(ms))
        return ms.value


class ExternalMemory:
    """Memory of another API on this GPU, imported from a file descriptor
    (hl_external_import); ``map`` gives device pointers into it for the *_d
    arguments (the reference's frame slots are wgpu buffers,
    crates/libhelio/src/frame.rs:228-234).  The caller keeps its descriptor."""

    def __init__(self, ctx: Context, fd: int, size: int):
        self.ctx = ctx
        self.size = int(size)
        h = C.c_void_p()
        ctx.lib.call("hl_external_import", ctx.h, int(fd), self.size, abi.HL_EXTERNAL_FD, C.byref(h))
        self.h = h.value

    def map(self, offset: int, size: int) -> int:
        p = C.c_void_p()
        self.ctx.lib.call("hl_external_map", self.h, int(offset), int(size), C.byref(p), ctx=self.ctx.h)
        return p.value

    def release(self) -> None:
        if self.h:
            h, self.h = self.h, None
            self.ctx.lib.call("hl_external_release", h, ctx=self.ctx.h)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.release()


@dataclass
class DeviceScene:
    """A HostScene resident in HBM."""
    host: object
    camera: DeviceArray
    lights: DeviceArray
    planes: dict
    gbuffer: abi.hl_gbuffer

    @property
    def num_lights(self) -> int:
        return len(self.host.lights)


def upload_scene(ctx: Context, scene) -> DeviceScene:
    cam = ctx.upload(np.frombuffer(bytes(scene.camera), dtype=np.uint8))
    lights = ctx.upload(scene.lights) if len(scene.lights) else ctx.alloc(128)
    planes = {k: ctx.upload(getattr(scene, k)) for k in ("albedo", "normal", "orm", "emissive", "depth")}
    g = abi.hl_gbuffer()
    g.width, g.height = scene.width, scene.height
    fa, ta, fn, tn = scene.formats
    for name, fmt, tb in (("albedo", fa, ta), ("orm", fa, ta), ("normal", fn, tn), ("emissive", fn, tn)):
        t = getattr(g, name)
        t.data_d, t.pitch_bytes, t.format = planes[name].ptr, scene.width * tb, fmt
    g.depth.data_d, g.depth.pitch_bytes, g.depth.format = planes["depth"].ptr, scene.width * 4, abi.HL_FMT_R32_FLOAT
    return DeviceScene(scene, cam, lights, planes, g)


def surface_desc(width: int, height: int, albedo, normal, orm, emissive, specular_f0=None) -> abi.hl_surface:
    """hl_surface over RGBA32F planes of width x height (DeviceArray or host numpy arrays)."""
    def ptr(p):
        if p is None:
            return None
        return p.ptr if isinstance(p, DeviceArray) else p.ctypes.data
    s = abi.hl_surface()
    s.width, s.height, s.pitch_bytes = width, height, 16 * width
    s.albedo_d, s.normal_d, s.orm_d, s.emissive_d = ptr(albedo), ptr(normal), ptr(orm), ptr(emissive)
    s.specular_f0_d = ptr(specular_f0)
    return s


def pack_gbuffer(ctx: Context, surface: abi.hl_surface, gbuffer: abi.hl_gbuffer) -> None:
    """The G-buffer pass's target writes (hl_gbuffer_pack): surface attributes
    into the G-buffer textures, F0 resolved into their alpha channels."""
    ctx.lib.call("hl_gbuffer_pack", ctx.h, C.byref(surface), C.byref(gbuffer), ctx=ctx.h)


class DeviceShadows:
    """Shadow atlases and light matrices resident in HBM (hl_shadow_maps)."""

    def __init__(self, ctx: Context, host):
        self.atlas = ctx.upload(host.atlas)
        self.static_atlas = ctx.upload(host.static_atlas)
        self.matrices = ctx.upload(host.matrices)
        d = abi.hl_shadow_maps()
        d.atlas_d, d.static_atlas_d, d.matrices_d = self.atlas.ptr, self.static_atlas.ptr, self.matrices.ptr
        d.num_layers, d.atlas_size, d.config = host.layers, host.size, host.config
        self.desc = d


class DeviceTexture:
    """An RGBA16F / RGBA32F image resident in HBM as an hl_texture (rows tightly packed)."""

    def __init__(self, ctx: Context, img: np.ndarray):
        self.buf = ctx.upload(img)
        t = abi.hl_texture()
        t.data_d, t.pitch_bytes = self.buf.ptr, img.shape[1] * img.itemsize * img.shape[2]
        t.format = abi.HL_FMT_RGBA16_FLOAT if img.dtype == np.float16 else abi.HL_FMT_RGBA32_FLOAT
        self.desc = t


class DeviceReflections:
    """hl_deferred_reflection's inputs resident in HBM (from a synth.HostReflections)."""

    def __init__(self, ctx: Context, host):
        self.host = host
        self.cubes = ctx.upload(host.cubes)
        self.caps = ctx.upload(host.captures) if len(host.captures) else None
        self.tex = {}
        r = abi.hl_reflection_inputs()
        r.env = host.env_host()
        r.env.data_d = self.cubes.ptr
        r.captures_d = self.caps.ptr if self.caps is not None else None
        for name, img in (("ssr", host.ssr), ("planar", host.planar), ("rc_cascade0", host.rc)):
            if img is not None:
                self.tex[name] = DeviceTexture(ctx, img)
                setattr(r, name, self.tex[name].desc)
        self.inputs = r


class ReflectionPass:
    """The deferred pass's second draw (fs_reflection, deferred_lighting.wgsl:1266-1341)
    over a frame that DeferredLightPass lit: adds the indirect specular composite."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def execute(self, scene: DeviceScene, refl: DeviceReflections, out: DeviceArray, flags: int = 0,
                globals_: abi.DeferredGlobals | None = None) -> None:
        g = globals_ if globals_ is not None else scene.host.globals
        self.ctx.lib.call("hl_deferred_reflection", self.ctx.h, scene.camera.ptr, C.byref(g),
                          C.byref(scene.gbuffer), C.byref(refl.inputs), out.ptr, scene.host.width * 16, flags,
                          ctx=self.ctx.h)


class TileGrid:
    """tile_light_lists / tile_light_counts (crates/libhelio/src/frame.rs:230-234)."""

    def __init__(self, ctx: Context, width: int, height: int):
        self.ntx = (width + 15) // 16
        self.nty = (height + 15) // 16
        self.lists = ctx.alloc(self.ntx * self.nty * abi.HL_MAX_LIGHTS_PER_TILE * 4)
        self.counts = ctx.alloc(self.ntx * self.nty * 4)

    def grid(self) -> abi.hl_light_grid:
        g = abi.hl_light_grid()
        g.mode, g.num_tiles_x, g.num_tiles_y = abi.HL_GRID_TILES, self.ntx, self.nty
        g.lists_d, g.counts_d = self.lists.ptr, self.counts.ptr
        g.lists_capacity = self.ntx * self.nty * abi.HL_MAX_LIGHTS_PER_TILE
        return g


class ClusterGrid:
    """Compact clustered lists: offsets / counts / pool."""

    def __init__(self, ctx: Context, params: abi.hl_cluster_params, pool_cap: int):
        self.params = params
        self.ncell = params.grid_x * params.grid_y * params.grid_z
        self.offsets = ctx.alloc(self.ncell * 4)
        self.counts = ctx.alloc(self.ncell * 4)
        self.pool_cap = int(pool_cap)
        self.pool = ctx.alloc(max(self.pool_cap, 1) * 4)
        self.overflow = ctx.alloc(16)

    def grid(self) -> abi.hl_light_grid:
        g = abi.hl_light_grid()
        g.mode = abi.HL_GRID_CLUSTERS
        g.lists_d, g.counts_d, g.offsets_d = self.pool.ptr, self.counts.ptr, self.offsets.ptr
        g.clusters = C.pointer(self.params)
        g.lists_capacity = self.pool_cap
        return g

    def download(self):
        off = self.offsets.download(np.uint32)[: self.ncell]
        cnt = self.counts.download(np.uint32)[: self.ncell]
        pool = self.pool.download(np.uint32)[: self.pool_cap]
        return off, cnt, pool


def light_cull_params(scene) -> abi.LightCullParams:
    """LightCullPass::prepare (helio-pass-light-cull/src/lib.rs:264-282)."""
    p = abi.LightCullParams()
    p.num_tiles_x = (scene.width + 15) // 16
    p.num_tiles_y = (scene.height + 15) // 16
    p.num_lights = len(scene.lights)
    p.screen_width, p.screen_height = scene.width, scene.height
    return p


class LightCullPass:
    """Tile (reference) or cluster light assignment behind the RenderPass shape."""

    def __init__(self, ctx: Context, width: int, height: int, clusters: abi.hl_cluster_params | None = None,
                 pool_cap: int = 0):
        self.ctx = ctx
        self.width, self.height = width, height
        self.clusters = clusters
        self.tiles = TileGrid(ctx, width, height) if clusters is None else None
        self.cgrid = ClusterGrid(ctx, clusters, pool_cap) if clusters is not None else None
        self.params = None

    def name(self) -> str:
        return "LightCull"

    def writes(self) -> tuple:
        return ("tile_light_lists", "tile_light_counts")

    def prepare(self, scene: DeviceScene) -> None:
        self.params = light_cull_params(scene.host)

    def execute(self, scene: DeviceScene) -> None:
        L = self.ctx.lib
        if self.tiles is not None:
            L.call("hl_light_cull_tiles", self.ctx.h, scene.camera.ptr, C.byref(self.params),
                   scene.lights.ptr, self.tiles.lists.ptr, self.tiles.counts.ptr)
        else:
            g = self.cgrid
            L.call("hl_light_cull_clusters", self.ctx.h, scene.camera.ptr, C.byref(self.clusters),
                   scene.lights.ptr, scene.num_lights, g.offsets.ptr, g.counts.ptr, g.pool.ptr,
                   g.pool_cap, g.overflow.ptr)

    def publish(self) -> abi.hl_light_grid:
        return self.tiles.grid() if self.tiles is not None else self.cgrid.grid()

    def overflow(self) -> int:
        f = C.c_uint32()
        self.ctx.lib.call("hl_query_overflow", self.ctx.h, C.byref(f))
        return f.value


class DeferredLightPass:
    """out_format: the pre_aa target's format, RGBA32F (default) or RGBA16F
    (HL_SHADE_OUT_RGBA16F; forward shading only)."""

    def __init__(self, ctx: Context, width: int, height: int, out_format: int = abi.HL_FMT_RGBA32_FLOAT):
        if out_format not in (abi.HL_FMT_RGBA32_FLOAT, abi.HL_FMT_RGBA16_FLOAT):
            raise ValueError("pre_aa must be RGBA32F or RGBA16F")
        self.ctx = ctx
        self.width, self.height = width, height
        self.out_f16 = out_format == abi.HL_FMT_RGBA16_FLOAT
        self.pitch = width * (8 if self.out_f16 else 16)
        self.out = ctx.alloc(self.pitch * height)
        self.debug_mode = 0
        self.globals = None

    def name(self) -> str:
        return "DeferredLight"

    def writes(self) -> tuple:
        return ("pre_aa",)

    def prepare(self, scene: DeviceScene) -> None:
        g = abi.DeferredGlobals()
        C.memmove(C.byref(g), C.byref(scene.host.globals), C.sizeof(g))
        g.debug_mode = self.debug_mode
        g.num_tiles_x = (self.width + 15) // 16
        self.globals = g

    def execute(self, scene: DeviceScene, grid: abi.hl_light_grid, flags: int = 0,
                shadows: "DeviceShadows | None" = None, rc: "DeviceTexture | None" = None) -> None:
        if self.out_f16:
            flags |= abi.HL_SHADE_OUT_RGBA16F
        if rc is not None:
            self.ctx.lib.call("hl_deferred_forward_ex", self.ctx.h, scene.camera.ptr, C.byref(self.globals),
                              C.byref(scene.gbuffer), scene.lights.ptr, scene.num_lights, C.byref(grid),
                              C.byref(shadows.desc) if shadows is not None else None, C.byref(rc.desc),
                              self.out.ptr, self.pitch, flags)
        elif shadows is None:
            self.ctx.lib.call("hl_deferred_forward", self.ctx.h, scene.camera.ptr, C.byref(self.globals),
                              C.byref(scene.gbuffer), scene.lights.ptr, scene.num_lights, C.byref(grid),
                              self.out.ptr, self.pitch, flags)
        else:
            self.ctx.lib.call("hl_deferred_forward_shadowed", self.ctx.h, scene.camera.ptr, C.byref(self.globals),
                              C.byref(scene.gbuffer), scene.lights.ptr, scene.num_lights, C.byref(grid),
                              C.byref(shadows.desc), self.out.ptr, self.pitch, flags)

    def forward_backward_l2(self, scene: DeviceScene, grid: abi.hl_light_grid, target: DeviceArray,
                            grads: "AdjointBuffers") -> None:
        """Forward into self.out, then the L2 loss against target and its
        gradients (hl_deferred_forward_backward_l2)."""
        if self.out_f16:
            raise ValueError("the fused adjoint writes an RGBA32F output")
        self.ctx.lib.call("hl_deferred_forward_backward_l2", self.ctx.h, scene.camera.ptr, C.byref(self.globals),
                          C.byref(scene.gbuffer), scene.lights.ptr, scene.num_lights, C.byref(grid), target.ptr,
                          self.width * 16, self.out.ptr, self.width * 16, grads.loss.ptr, grads.d_pixel.ptr,
                          grads.d_pos.ptr, grads.d_int.ptr)

    def backward(self, scene: DeviceScene, grid: abi.hl_light_grid, d_out: DeviceArray,
                 grads: "AdjointBuffers") -> None:
        """Gradients for a given dL/dout (RGBA32F, alpha ignored), hl_deferred_backward."""
        self.ctx.lib.call("hl_deferred_backward", self.ctx.h, scene.camera.ptr, C.byref(self.globals),
                          C.byref(scene.gbuffer), scene.lights.ptr, scene.num_lights, C.byref(grid), d_out.ptr,
                          self.width * 16, grads.d_pixel.ptr, grads.d_pos.ptr, grads.d_int.ptr)

    def count_pairs(self, scene: DeviceScene, grid: abi.hl_light_grid) -> int:
        n = C.c_uint64()
        self.ctx.lib.call("hl_count_pixel_light_pairs", self.ctx.h, scene.camera.ptr, C.byref(self.globals),
                          C.byref(scene.gbuffer), scene.lights.ptr, scene.num_lights, C.byref(grid), C.byref(n))
        return n.value

    def walk_stats(self, scene: DeviceScene, grid: abi.hl_light_grid) -> tuple[int, int]:
        """(Q, wavefront-light visits of the forward walk), hl_shade_walk_stats."""
        n, v = C.c_uint64(), C.c_uint64()
        self.ctx.lib.call("hl_shade_walk_stats", self.ctx.h, scene.camera.ptr, C.byref(self.globals),
                          C.byref(scene.gbuffer), scene.lights.ptr, scene.num_lights, C.byref(grid), C.byref(n),
                          C.byref(v))
        return n.value, v.value

    def output(self) -> np.ndarray:
        """The target as float32 (RGBA16F values widened exactly)."""
        if self.out_f16:
            return self.out.download(np.float16, (self.height, self.width, 4)).astype(np.float32)
        return self.out.download(np.float32, (self.height, self.width, 4))


class ForwardLitPass:
    """ForwardLit's fragment stage (hl_forward_lit): fs_main of forward_lit.wgsl
    over a batch of fragments, lit from the light grid LightCull published --
    the second consumer of that grid
    (crates/passes/3d/helio-pass-forward-lit/src/lib.rs:328-615).  The
    rasteriser that produces the fragments and the blend into pre_aa stay with
    the caller; one RGBA value per fragment comes back."""

    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx = ctx
        self.width, self.height = width, height
        self.globals = None
        self.out = None

    def name(self) -> str:
        return "ForwardLit"

    def reads(self) -> tuple:
        return ("main_scene", "depth", "pre_aa", "cluster_light_grid")

    def writes(self) -> tuple:
        return ("pre_aa",)

    def prepare(self, scene: DeviceScene) -> None:
        from .synth import forward_globals
        self.globals = forward_globals(scene.host)

    def _args(self, scene, fragments, num_fragments, grid, lights, entity, transforms, num_transforms):
        lights = lights if lights is not None else scene.lights
        return (self.ctx.h, scene.camera.ptr, C.byref(self.globals), fragments.ptr if num_fragments else None,
                num_fragments, lights.ptr, scene.num_lights, entity.ptr if entity is not None else None,
                transforms.ptr if transforms is not None else None, num_transforms, C.byref(grid))

    def execute(self, scene: DeviceScene, fragments: DeviceArray, num_fragments: int, grid: abi.hl_light_grid,
                lights: DeviceArray | None = None, entity: DeviceArray | None = None,
                transforms: DeviceArray | None = None, num_transforms: int = 0) -> DeviceArray:
        """Shades num_fragments fragments (abi.FRAGMENT_DTYPE records) into
        self.out (RGBA32F per fragment).  lights defaults to the scene's."""
        if self.out is None or self.out.nbytes < 16 * max(num_fragments, 1):
            self.out = self.ctx.alloc(16 * max(num_fragments, 1))
        self.ctx.lib.call("hl_forward_lit", *self._args(scene, fragments, num_fragments, grid, lights, entity,
                                                        transforms, num_transforms),
                          self.out.ptr, 0, ctx=self.ctx.h)
        return self.out

    def count_pairs(self, scene: DeviceScene, fragments: DeviceArray, num_fragments: int,
                    grid: abi.hl_light_grid, lights: DeviceArray | None = None, entity: DeviceArray | None = None,
                    transforms: DeviceArray | None = None, num_transforms: int = 0) -> int:
        n = C.c_uint64()
        self.ctx.lib.call("hl_forward_lit_count_pairs", *self._args(scene, fragments, num_fragments, grid, lights,
                                                                    entity, transforms, num_transforms),
                          C.byref(n), ctx=self.ctx.h)
        return n.value

    def output(self, num_fragments: int) -> np.ndarray:
        return self.out.download(np.float32, (num_fragments, 4), nbytes=16 * num_fragments)


class AdjointBuffers:
    """Outputs of the adjoint: loss (FP64), 5 per-pixel gradient planes
    (albedo r, g, b, roughness, metallic) and per-light position / intensity
    gradients."""

    def __init__(self, ctx: Context, width: int, height: int, num_lights: int):
        n = max(num_lights, 1)
        self.loss = ctx.alloc(8)
        self.d_pixel = ctx.alloc(abi.HL_PIXEL_GRAD_PLANES * width * height * 4)
        self.d_pos = ctx.alloc(3 * n * 4)
        self.d_int = ctx.alloc(n * 4)
