"""Batched multi-view frames over the C ABI (hl_batch_*).

The reference renders each view with its own graph submit
(crates/helio/src/renderer/render.rs:804-873); here a BatchPlan holds every
view and replays one captured frame.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi, synth
from .device import Context, DeviceArray, DeviceScene, upload_scene


@dataclass
class BatchView:
    scene: DeviceScene
    subset: DeviceArray | None
    num_lights: int
    out: DeviceArray
    adjoint: bool = False
    target: DeviceArray | None = None
    loss: DeviceArray | None = None
    d_pixel: DeviceArray | None = None
    d_pos: DeviceArray | None = None
    d_int: DeviceArray | None = None
    clusters: abi.hl_cluster_params | None = None  # fixed grid; None = the plan chooses

    def desc(self) -> abi.hl_view_desc:
        d = abi.hl_view_desc()
        sc = self.scene
        d.camera_d = sc.camera.ptr
        C.memmove(C.byref(d.camera), C.byref(sc.host.camera), C.sizeof(abi.GpuCameraUniforms))
        C.memmove(C.byref(d.globals), C.byref(sc.host.globals), C.sizeof(abi.DeferredGlobals))
        d.gbuffer = sc.gbuffer
        d.light_indices_d = self.subset.ptr if self.subset is not None else None
        d.num_lights = self.num_lights
        d.out_d = self.out.ptr
        d.out_pitch_bytes = sc.host.width * 16
        if self.clusters is not None:
            d.clusters = C.pointer(self.clusters)
        if self.adjoint:
            d.target_d = self.target.ptr
            d.target_pitch_bytes = sc.host.width * 16
            d.loss_d = self.loss.ptr
            d.d_pixel_d = self.d_pixel.ptr
            d.d_light_pos_d = self.d_pos.ptr
            d.d_light_intensity_d = self.d_int.ptr
        return d


class BatchPlan:
    def __init__(self, ctx: Context, pool: DeviceArray, pool_size: int, views: list[BatchView],
                 autotune: bool = False, max_trials: int = 0, cache_path: str | None = None):
        self.ctx = ctx
        self.views = views
        descs = (abi.hl_view_desc * len(views))(*[v.desc() for v in views])
        opts = abi.hl_autotune_opts()
        opts.enable = 1 if autotune else 0
        opts.max_trials = max_trials
        opts.cache_path = cache_path.encode() if cache_path else None
        h = C.c_void_p()
        ctx.lib.call("hl_batch_plan_create", ctx.h, pool.ptr, pool_size, descs, len(views), C.byref(opts),
                     C.byref(h), ctx=ctx.h)
        self.h = h.value

    def capture(self) -> None:
        self.ctx.lib.call("hl_batch_capture", self.h, ctx=self.ctx.h)

    def launch(self) -> None:
        self.ctx.lib.call("hl_batch_launch", self.h, ctx=self.ctx.h)

    def config(self, v: int) -> abi.hl_view_config:
        c = abi.hl_view_config()
        self.ctx.lib.call("hl_batch_get_config", self.h, v, C.byref(c), ctx=self.ctx.h)
        return c

    def destroy(self) -> None:
        if self.h:
            self.ctx.lib.call("hl_plan_destroy", self.h, ctx=self.ctx.h)
            self.h = None


def make_multiview(ctx: Context, small: bool = False, adjoint_eyes: bool = True):
    """The 50-view workload (48 probe faces + 2 XR eyes, 8192-light pool);
    small=True shrinks every view 8x per side.  Returns (pool, pool_size,
    views, host scenes, pool lights)."""
    specs, subsets, pool_spec = synth.multiview_specs(small=small)
    pool_lights = synth.make_lights(pool_spec)
    pool = ctx.upload(pool_lights)
    views, hosts = [], []
    for i, (s, sub) in enumerate(zip(specs, subsets)):
        sc = synth.make_scene(s)
        sc.lights = pool_lights[sub]  # host copy of the view's lights (for checks)
        ds = upload_scene(ctx, sc)
        out = ctx.alloc(s.width * s.height * 16)
        adj = adjoint_eyes and s.name.startswith("eye")
        v = BatchView(scene=ds, subset=ctx.upload(sub.astype(np.uint32)), num_lights=len(sub), out=out,
                      adjoint=adj)
        if adj:
            v.target = ctx.alloc(s.width * s.height * 16)
            v.loss = ctx.alloc(8)
            v.d_pixel = ctx.alloc(5 * s.width * s.height * 4)
            v.d_pos = ctx.alloc(3 * len(pool_lights) * 4)
            v.d_int = ctx.alloc(len(pool_lights) * 4)
        views.append(v)
        hosts.append((sc, sub))
    return pool, len(pool_lights), views, hosts, pool_lights
