"""ctypes binding of oracle/liboracle.so (the CPU checker; tests only)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from helio_amd import abi

ROOT = Path(__file__).resolve().parent.parent
P = C.POINTER
vp = C.c_void_p
u32 = C.c_uint32


class or64_pixel(C.Structure):
    _fields_ = [("wp", C.c_double * 3), ("N", C.c_double * 3), ("V", C.c_double * 3),
                ("albedo", C.c_double * 3), ("rough", C.c_double), ("metal", C.c_double),
                ("ao", C.c_double), ("emissive", C.c_double * 3), ("alpha", C.c_double),
                ("sky", C.c_int)]


_SIGS = {
    "or_light_cull_tiles": (None, [P(abi.GpuCameraUniforms), P(abi.LightCullParams), vp, vp, vp]),
    "or_light_cull_clusters": (u32, [P(abi.GpuCameraUniforms), P(abi.hl_cluster_params), vp, u32, vp, vp, vp, u32]),
    "or_cluster_hits": (u32, [P(abi.GpuCameraUniforms), P(abi.hl_cluster_params), vp, u32, u32]),
    "or_slice_of_depth": (u32, [P(abi.hl_cluster_params), C.c_float]),
    "or_gbuffer_pack": (None, [P(abi.hl_surface), P(abi.hl_gbuffer)]),
    "or_float_to_half": (C.c_uint16, [C.c_float]),
    "or_unorm8_encode": (u32, [C.c_float]),
    "or_f0_metallic": (C.c_float, [C.c_float, C.c_float]),
    "or_forward_lit": (C.c_uint64, [P(abi.GpuCameraUniforms), P(abi.ForwardLitGlobals), vp, u32, vp, u32, vp, vp,
                                    u32, P(abi.hl_light_grid), vp]),
    "or_light_direction_from_rotation": (None, [P(C.c_float), P(C.c_float)]),
    "or_forward_pbr_direct_light": (None, [vp, P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float),
                                           P(C.c_float), P(C.c_float), C.c_float, C.c_float, P(C.c_float)]),
    "or_world_pos": (None, [P(abi.GpuCameraUniforms), u32, u32, u32, u32, C.c_float, P(C.c_float)]),
    "or_shade_ex": (C.c_uint64, [P(abi.GpuCameraUniforms), P(abi.DeferredGlobals), P(abi.hl_gbuffer), vp, u32,
                                 P(abi.hl_light_grid), P(abi.hl_shadow_maps), P(abi.hl_texture), vp, u32, u32, u32,
                                 u32]),
    "or_rc_oct_decode": (None, [C.c_float, C.c_float, P(C.c_float)]),
    "or_rc_irradiance": (None, [P(abi.hl_texture), P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float),
                                P(C.c_float)]),
    "or_shade_shadowed": (C.c_uint64, [P(abi.GpuCameraUniforms), P(abi.DeferredGlobals), P(abi.hl_gbuffer), vp, u32,
                                       P(abi.hl_light_grid), P(abi.hl_shadow_maps), vp, u32, u32, u32, u32]),
    "or_shadow_factor": (C.c_float, [P(abi.hl_shadow_maps), vp, u32, P(C.c_float), P(C.c_float), P(C.c_float),
                                     P(C.c_float), C.c_float, C.c_float]),
    "or_sample_cascade_shadow": (C.c_float, [P(abi.hl_shadow_maps), u32, u32, C.c_float, P(C.c_float), C.c_float,
                                             C.c_float]),
    "or_sample_cascade_shadow_pcss": (C.c_float, [P(abi.hl_shadow_maps), u32, u32, P(C.c_float), C.c_float,
                                                  C.c_float]),
    "or_sample_compare": (C.c_float, [vp, u32, u32, C.c_float, C.c_float, C.c_float]),
    "or_sincos": (None, [C.c_float, P(C.c_float), P(C.c_float)]),
    "or_hash22": (C.c_float, [C.c_float, C.c_float]),
    "or_vogel": (None, [u32, u32, C.c_float, P(C.c_float)]),
    "or_point_light_face": (u32, [P(C.c_float)]),
    "or_shade": (C.c_uint64, [P(abi.GpuCameraUniforms), P(abi.DeferredGlobals), P(abi.hl_gbuffer), vp, u32,
                              P(abi.hl_light_grid), vp, u32, u32, u32, u32]),
    "or_count_pairs": (C.c_uint64, [P(abi.GpuCameraUniforms), P(abi.DeferredGlobals), P(abi.hl_gbuffer), vp, u32,
                                    P(abi.hl_light_grid)]),
    "or_check_conservative": (C.c_uint64, [P(abi.GpuCameraUniforms), P(abi.DeferredGlobals), P(abi.hl_gbuffer),
                                           vp, u32, P(abi.hl_light_grid), u32, u32, P(C.c_uint64)]),
    "or_pbr_direct_light": (None, [vp, P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float),
                                   C.c_float, C.c_float, P(C.c_float)]),
    "or_cube_bytes": (C.c_size_t, [u32, u32, u32]),
    "or_cube_face": (None, [P(C.c_float), P(u32), P(C.c_float), P(C.c_float)]),
    "or_cube_sample": (None, [P(abi.hl_cube_array), u32, P(C.c_float), C.c_float, P(C.c_float)]),
    "or_capture_weight": (C.c_float, [P(C.c_float), P(abi.GpuReflectionCapture)]),
    "or_capture_parallax": (None, [P(C.c_float), P(C.c_float), P(abi.GpuReflectionCapture), P(C.c_float)]),
    "or_reflection_env": (None, [P(abi.DeferredGlobals), P(abi.hl_reflection_inputs), P(C.c_float), P(C.c_float),
                                 C.c_float, P(C.c_float)]),
    "or_env_brdf_approx": (None, [C.c_float, C.c_float, P(C.c_float)]),
    "or_reflection": (C.c_uint64, [P(abi.GpuCameraUniforms), P(abi.DeferredGlobals), P(abi.hl_gbuffer),
                                   P(abi.hl_reflection_inputs), vp, u32, u32, u32, u32]),
    "or64_decode": (None, [P(abi.GpuCameraUniforms), P(abi.hl_gbuffer), vp]),
    "or64_shade_pixel": (C.c_uint64, [P(or64_pixel), P(abi.DeferredGlobals), vp, u32, P(C.c_double)]),
    "or64_forward": (C.c_double, [vp, u32, u32, P(abi.DeferredGlobals), vp, u32, vp, vp, u32]),
    "or64_backward_l2": (None, [vp, u32, u32, P(abi.DeferredGlobals), vp, u32, vp, u32, vp, vp, vp]),
    "or64_backward": (None, [vp, u32, u32, P(abi.DeferredGlobals), vp, u32, vp, u32, vp, vp, vp]),
    "or64_fd": (C.c_double, [vp, u32, u32, P(abi.DeferredGlobals), vp, u32, vp, u32, C.c_int, u32,
                             C.c_double, P(C.c_int)]),
}

_LIB = None


def oracle() -> C.CDLL:
    global _LIB
    if _LIB is None:
        p = ROOT / "oracle" / "liboracle.so"
        if not p.exists():
            raise FileNotFoundError(f"{p} missing: run `python -m helio_amd.build`")
        lib = C.CDLL(str(p))
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _LIB = lib
    return _LIB


def f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def cull_tiles(scene):
    from helio_amd.device import light_cull_params
    p = light_cull_params(scene)
    tiles = p.num_tiles_x * p.num_tiles_y
    lists = np.zeros(tiles * 64, dtype=np.uint32)
    counts = np.zeros(tiles, dtype=np.uint32)
    oracle().or_light_cull_tiles(C.byref(scene.camera), C.byref(p), scene.lights.ctypes.data,
                                 lists.ctypes.data, counts.ctypes.data)
    return lists.reshape(tiles, 64), counts


def cull_clusters(scene, params, pool_cap):
    n = params.grid_x * params.grid_y * params.grid_z
    off = np.zeros(n, dtype=np.uint32)
    cnt = np.zeros(n, dtype=np.uint32)
    pool = np.zeros(max(pool_cap, 1), dtype=np.uint32)
    flags = oracle().or_light_cull_clusters(C.byref(scene.camera), C.byref(params),
                                            scene.lights.ctypes.data if len(scene.lights) else None,
                                            len(scene.lights), off.ctypes.data, cnt.ctypes.data,
                                            pool.ctypes.data, pool_cap)
    return off, cnt, pool[:pool_cap], flags


def tile_grid_host(scene, lists, counts):
    g = abi.hl_light_grid()
    g.mode = abi.HL_GRID_TILES
    g.num_tiles_x = (scene.width + 15) // 16
    g.num_tiles_y = (scene.height + 15) // 16
    g.lists_d, g.counts_d = lists.ctypes.data, counts.ctypes.data
    g.lists_capacity = lists.size
    return g


def cluster_grid_host(params, off, cnt, pool):
    g = abi.hl_light_grid()
    g.mode = abi.HL_GRID_CLUSTERS
    g.lists_d, g.counts_d, g.offsets_d = pool.ctypes.data, cnt.ctypes.data, off.ctypes.data
    g.clusters = C.pointer(params)
    g.lists_capacity = pool.size
    g._arrays = (off, cnt, pool)  # the struct holds raw addresses: keep the arrays alive with it
    return g


def shade(scene, grid, globals_=None, flags=0, rows=None, out=None, shadows=None, rc=None):
    """FP32 restatement of fs_main; returns (rgba [H,W,4], pairs).  shadows:
    a synth.HostShadows to multiply the terms of shadow-casting lights by
    shadow_factor; rc: an rc_cascade0 probe atlas (synth.make_rc_probes) for the
    radiance-cascade indirect diffuse (used when the globals have has_rc_gi)."""
    W, H = scene.width, scene.height
    if out is None:
        out = np.zeros((H, W, 4), dtype=np.float32)
    g = globals_ if globals_ is not None else scene.globals
    gb = scene.gbuffer_host()
    r0, r1 = rows if rows else (0, 0)
    sh = C.byref(shadows.host()) if shadows is not None else None
    rt = C.byref(rc_texture_of(rc)) if rc is not None else None
    pairs = oracle().or_shade_ex(C.byref(scene.camera), C.byref(g), C.byref(gb),
                                 scene.lights.ctypes.data if len(scene.lights) else None, len(scene.lights),
                                 C.byref(grid), sh, rt, out.ctypes.data, W * 16, flags, r0, r1)
    return out, pairs


def reflection(scene, refl, out=None, flags=0, rows=None, globals_=None):
    """FP32 restatement of fs_reflection over the frame out (added to it, or
    stored with abi.HL_REFLECTION_STORE); refl: a synth.HostReflections.
    Returns (out, pixels written)."""
    W, H = scene.width, scene.height
    if out is None:
        out = np.zeros((H, W, 4), dtype=np.float32)
    g = globals_ if globals_ is not None else scene.globals
    gb = scene.gbuffer_host()
    r0, r1 = rows if rows else (0, H)
    inp = refl.inputs_host()
    n = oracle().or_reflection(C.byref(scene.camera), C.byref(g), C.byref(gb), C.byref(inp), out.ctypes.data,
                               W * 16, flags, r0, r1)
    return out, n


def rc_texture_of(atlas):
    from helio_amd.synth import rc_texture_host
    return rc_texture_host(atlas)


def forward_lit(camera, globals_, frags, lights, grid, entity=None, transforms=None, count_only=False):
    """FP32 restatement of forward_lit.wgsl fs_main; returns (rgba [n, 4] or None, pairs)."""
    n = len(frags)
    out = None if count_only else np.zeros((n, 4), dtype=np.float32)
    pairs = oracle().or_forward_lit(C.byref(camera), C.byref(globals_), frags.ctypes.data if n else None, n,
                                    lights.ctypes.data if len(lights) else None, len(lights),
                                    entity.ctypes.data if entity is not None else None,
                                    transforms.ctypes.data if transforms is not None else None,
                                    len(transforms) if transforms is not None else 0, C.byref(grid),
                                    out.ctypes.data if out is not None else None)
    return out, pairs


def all_lights_grid(num_lights, width, height):
    """One cluster covering the screen that lists every light in order (a grid
    that cannot miss a light)."""
    p = abi.hl_cluster_params()
    p.grid_x = p.grid_y = p.grid_z = 1
    p.cell_w, p.cell_h, p.width, p.height = width, height, width, height
    p.near_plane, p.far_plane = 0.1, 1000.0
    off = np.zeros(1, np.uint32)
    cnt = np.array([num_lights], np.uint32)
    pool = np.arange(max(num_lights, 1), dtype=np.uint32)
    g = cluster_grid_host(p, off, cnt, pool)
    g._params = p
    g.clusters = C.pointer(p)
    return g
