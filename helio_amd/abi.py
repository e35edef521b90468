"""ctypes view of the C ABI in include/helio_hip/helio_hip.h.

The struct layouts here are the same bytes as the C headers (and therefore as
libhelio's Rust records); ``check_layouts`` asserts the sizes.  Functions of
libhelio_hip.so return an ``hl_status``; :class:`Lib` raises
:class:`HelioError` for any non-zero status, carrying ``hl_last_error``.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent

HL_OK, HL_EINVAL, HL_EHIP, HL_ECAPACITY, HL_EUNSUPPORTED, HL_ENOMEM = 0, -1, -2, -3, -4, -5
HL_LIGHT_DIRECTIONAL, HL_LIGHT_POINT, HL_LIGHT_SPOT, HL_LIGHT_AREA = 0, 1, 2, 3
HL_NO_SHADOW = 0xFFFFFFFF
HL_TILE_SIZE, HL_MAX_LIGHTS_PER_TILE = 16, 64
HL_FMT_RGBA8_UNORM, HL_FMT_RGBA16_FLOAT, HL_FMT_RGBA32_FLOAT, HL_FMT_R32_FLOAT, HL_FMT_RG16_FLOAT = 1, 2, 3, 4, 5
HL_GRID_TILES, HL_GRID_CLUSTERS = 1, 2
HL_MAX_SLICES = 64
HL_SHADE_CLEAR_SKY = 1
HL_SHADE_LIGHTS_FROM_CULL = 2
HL_SHADE_OUT_RGBA16F = 4
HL_OVERFLOW_CAP, HL_OVERFLOW_POOL = 1, 2
HL_PIXEL_GRAD_PLANES = 5
HL_BATCH_MAX_VIEWS = 255
HL_NOALLOC_MAX_LIGHTS = 65536
HL_NOALLOC_SCRATCH_BYTES = 16384 + 80 * HL_NOALLOC_MAX_LIGHTS
HL_EXTERNAL_FD, HL_EXTERNAL_MAP_ALIGN = 1, 256

# GpuLight (crates/libhelio/src/light.rs:60-106), 128 bytes.
GPU_LIGHT_DTYPE = np.dtype([
    ("position_range", "<f4", 4), ("direction_outer", "<f4", 4), ("color_intensity", "<f4", 4),
    ("shadow_index", "<u4"), ("light_type", "<u4"), ("inner_angle", "<f4"), ("_pad", "<u4"),
    ("god_rays_enabled", "<u4"), ("god_rays_density", "<f4"), ("god_rays_weight", "<f4"),
    ("god_rays_decay", "<f4"), ("god_rays_exposure", "<f4"), ("flare_enabled", "<u4"),
    ("flare_type", "<u4"), ("flare_intensity", "<f4"), ("flare_scale", "<f4"),
    ("flare_tint_r", "<f4"), ("flare_tint_g", "<f4"), ("flare_tint_b", "<f4"),
    ("ies_profile_index", "<i4"), ("light_function_index", "<i4"), ("ies_angle_scale", "<f4"),
    ("ies_angle_offset", "<f4"),
])


class GpuCameraUniforms(C.Structure):
    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("view_proj", C.c_float * 16),
                ("inv_view_proj", C.c_float * 16), ("position_near", C.c_float * 4),
                ("forward_far", C.c_float * 4), ("jitter_frame", C.c_float * 4),
                ("prev_view_proj", C.c_float * 16)]


class LightCullParams(C.Structure):
    _fields_ = [("num_tiles_x", C.c_uint32), ("num_tiles_y", C.c_uint32), ("num_lights", C.c_uint32),
                ("screen_width", C.c_uint32), ("screen_height", C.c_uint32),
                ("_pad0", C.c_uint32), ("_pad1", C.c_uint32), ("_pad2", C.c_uint32)]


class DeferredGlobals(C.Structure):
    _fields_ = [("frame", C.c_uint32), ("delta_time", C.c_float), ("light_count", C.c_uint32),
                ("ambient_intensity", C.c_float), ("ambient_color", C.c_float * 4),
                ("rc_world_min", C.c_float * 4), ("rc_world_max", C.c_float * 4),
                ("csm_splits", C.c_float * 4), ("debug_mode", C.c_uint32), ("has_rc_gi", C.c_uint32),
                ("num_tiles_x", C.c_uint32), ("reflection_capture_count", C.c_uint32),
                ("enable_reflections", C.c_uint32), ("enable_env_reflections", C.c_uint32),
                ("_pad", C.c_uint32 * 2)]


class GpuMaterial(C.Structure):
    _fields_ = [("base_color", C.c_float * 4), ("emissive", C.c_float * 4),
                ("roughness_metallic", C.c_float * 4), ("tex_base_color", C.c_uint32),
                ("tex_normal", C.c_uint32), ("tex_roughness", C.c_uint32),
                ("tex_emissive", C.c_uint32), ("tex_occlusion", C.c_uint32), ("workflow", C.c_uint32),
                ("flags", C.c_uint32), ("material_class", C.c_uint32), ("class_params", C.c_float * 4)]


class ForwardLitGlobals(C.Structure):
    _fields_ = [("frame", C.c_uint32), ("delta_time", C.c_float), ("light_count", C.c_uint32),
                ("ambient_intensity", C.c_float), ("ambient_color", C.c_float * 4),
                ("num_tiles_x", C.c_uint32), ("num_tiles_y", C.c_uint32), ("screen_width", C.c_float),
                ("screen_height", C.c_float)]


class hl_transform(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("rotation", C.c_float * 3), ("scale", C.c_float * 3)]


class hl_fragment(C.Structure):
    _fields_ = [("world_pos", C.c_float * 3), ("pixel", C.c_uint32), ("normal", C.c_float * 3),
                ("alpha", C.c_float), ("albedo", C.c_float * 3), ("ao", C.c_float),
                ("specular_f0", C.c_float * 3), ("roughness", C.c_float), ("emissive", C.c_float * 3),
                ("metallic", C.c_float)]


# numpy views of the same records (fragments and transforms are built as arrays)
FRAGMENT_DTYPE = np.dtype([("world_pos", "<f4", 3), ("pixel", "<u4"), ("normal", "<f4", 3), ("alpha", "<f4"),
                           ("albedo", "<f4", 3), ("ao", "<f4"), ("specular_f0", "<f4", 3),
                           ("roughness", "<f4"), ("emissive", "<f4", 3), ("metallic", "<f4")])
TRANSFORM_DTYPE = np.dtype([("position", "<f4", 3), ("rotation", "<f4", 3), ("scale", "<f4", 3)])


class CascadeConfig(C.Structure):
    _fields_ = [("split_distance", C.c_float), ("depth_bias", C.c_float), ("filter_radius", C.c_float),
                ("pcss_light_size", C.c_float)]


class ShadowConfig(C.Structure):
    _fields_ = [("cascades", CascadeConfig * 4), ("enable_pcss", C.c_uint32), ("pcss_blocker_samples", C.c_uint32),
                ("pcss_filter_samples", C.c_uint32), ("pcf_sample_count", C.c_uint32)]


class hl_shadow_maps(C.Structure):
    _fields_ = [("atlas_d", C.c_void_p), ("static_atlas_d", C.c_void_p), ("matrices_d", C.c_void_p),
                ("num_layers", C.c_uint32), ("atlas_size", C.c_uint32), ("config", ShadowConfig)]


class hl_texture(C.Structure):
    _fields_ = [("data_d", C.c_void_p), ("pitch_bytes", C.c_uint32), ("format", C.c_uint32)]


HL_CAPTURE_SHAPE_SPHERE, HL_CAPTURE_SHAPE_BOX = 0, 1
HL_REFLECTION_STORE = 1


class GpuReflectionCapture(C.Structure):
    _fields_ = [("position_radius", C.c_float * 4), ("extents_transition", C.c_float * 4),
                ("world_to_local", C.c_float * 16), ("cubemap_index", C.c_int32), ("shape", C.c_uint32),
                ("mobility", C.c_uint32), ("brightness", C.c_float)]


CAPTURE_DTYPE = np.dtype([("position_radius", "<f4", 4), ("extents_transition", "<f4", 4),
                          ("world_to_local", "<f4", 16), ("cubemap_index", "<i4"), ("shape", "<u4"),
                          ("mobility", "<u4"), ("brightness", "<f4")])


class hl_cube_array(C.Structure):
    _fields_ = [("data_d", C.c_void_p), ("face_size", C.c_uint32), ("mip_levels", C.c_uint32),
                ("num_cubes", C.c_uint32), ("format", C.c_uint32)]


class hl_reflection_inputs(C.Structure):
    _fields_ = [("env", hl_cube_array), ("captures_d", C.c_void_p), ("ssr", hl_texture), ("planar", hl_texture),
                ("rc_cascade0", hl_texture)]


class hl_gbuffer(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("albedo", hl_texture),
                ("normal", hl_texture), ("orm", hl_texture), ("emissive", hl_texture),
                ("depth", hl_texture), ("lightmap_uv", hl_texture)]


class hl_cluster_params(C.Structure):
    _fields_ = [("grid_x", C.c_uint32), ("grid_y", C.c_uint32), ("grid_z", C.c_uint32),
                ("cell_w", C.c_uint32), ("cell_h", C.c_uint32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("cap", C.c_uint32), ("near_plane", C.c_float),
                ("far_plane", C.c_float), ("r_pad_column", C.c_float), ("reserved", C.c_float),
                ("depth_split", C.c_float * (HL_MAX_SLICES + 1)),
                ("slice_near", C.c_float * HL_MAX_SLICES), ("slice_far", C.c_float * HL_MAX_SLICES),
                ("slice_r_pad", C.c_float * HL_MAX_SLICES)]


class hl_surface(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("pitch_bytes", C.c_uint32),
                ("reserved", C.c_uint32), ("albedo_d", C.c_void_p), ("normal_d", C.c_void_p),
                ("orm_d", C.c_void_p), ("emissive_d", C.c_void_p), ("specular_f0_d", C.c_void_p)]


class hl_light_grid(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("num_tiles_x", C.c_uint32), ("num_tiles_y", C.c_uint32),
                ("lists_d", C.c_void_p), ("counts_d", C.c_void_p), ("offsets_d", C.c_void_p),
                ("clusters", C.POINTER(hl_cluster_params)), ("lists_capacity", C.c_uint32),
                ("reserved", C.c_uint32)]


class hl_view_desc(C.Structure):
    _fields_ = [("camera_d", C.c_void_p), ("camera", GpuCameraUniforms), ("globals", DeferredGlobals),
                ("gbuffer", hl_gbuffer), ("light_indices_d", C.c_void_p), ("num_lights", C.c_uint32),
                ("out_d", C.c_void_p), ("out_pitch_bytes", C.c_uint32), ("target_d", C.c_void_p),
                ("target_pitch_bytes", C.c_uint32), ("loss_d", C.c_void_p), ("d_pixel_d", C.c_void_p),
                ("d_light_pos_d", C.c_void_p), ("d_light_intensity_d", C.c_void_p),
                ("clusters", C.POINTER(hl_cluster_params))]


class hl_autotune_opts(C.Structure):
    _fields_ = [("enable", C.c_uint32), ("max_trials", C.c_uint32), ("cache_path", C.c_char_p)]


class hl_view_config(C.Structure):
    _fields_ = [("wave_tile_w", C.c_uint32), ("wave_tile_h", C.c_uint32), ("grid_x", C.c_uint32),
                ("grid_y", C.c_uint32), ("grid_z", C.c_uint32), ("est_ms", C.c_float)]


def check_layouts() -> None:
    assert GPU_LIGHT_DTYPE.itemsize == 128
    assert C.sizeof(GpuCameraUniforms) == 368
    assert C.sizeof(LightCullParams) == 32
    assert C.sizeof(DeferredGlobals) == 112
    assert C.sizeof(GpuMaterial) == 96
    assert C.sizeof(ForwardLitGlobals) == 48
    assert C.sizeof(hl_transform) == 36 == TRANSFORM_DTYPE.itemsize
    assert C.sizeof(hl_fragment) == 80 == FRAGMENT_DTYPE.itemsize
    assert C.sizeof(ShadowConfig) == 80
    assert C.sizeof(GpuReflectionCapture) == 112 == CAPTURE_DTYPE.itemsize
    assert C.sizeof(hl_cube_array) == 24


check_layouts()


class HelioError(RuntimeError):
    """A non-zero hl_status (the reference's helio_core::Error counterpart)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"hl_status {status}: {message}")
        self.status = status


vp = C.c_void_p
u32 = C.c_uint32
P = C.POINTER

# name -> (restype, argtypes); restype None means hl_status
SIGNATURES: dict[str, tuple] = {
    "hl_version": (C.c_char_p, []),
    "hl_status_string": (C.c_char_p, [C.c_int32]),
    "hl_context_create": (None, [C.c_int, P(vp)]),
    "hl_context_destroy": (None, [vp]),
    "hl_set_stream": (None, [vp, vp]),
    "hl_get_stream": (vp, [vp]),
    "hl_synchronize": (None, [vp]),
    "hl_last_error": (C.c_char_p, [vp]),
    "hl_alloc": (None, [vp, C.c_size_t, P(vp)]),
    "hl_free": (None, [vp, vp]),
    "hl_memcpy_htod": (None, [vp, vp, vp, C.c_size_t]),
    "hl_memcpy_dtoh": (None, [vp, vp, vp, C.c_size_t]),
    "hl_memcpy_dtod": (None, [vp, vp, vp, C.c_size_t]),
    "hl_memset_d32": (None, [vp, vp, u32, C.c_size_t]),
    "hl_event_create": (None, [vp, P(vp)]),
    "hl_event_destroy": (None, [vp]),
    "hl_event_record": (None, [vp, vp]),
    "hl_event_elapsed_ms": (None, [vp, vp, P(C.c_float)]),
    "hl_cluster_params_init": (None, [P(GpuCameraUniforms), u32, u32, u32, u32, u32, u32,
                                      P(hl_cluster_params)]),
    "hl_light_cull_tiles": (None, [vp, vp, P(LightCullParams), vp, vp, vp]),
    "hl_light_cull_clusters": (None, [vp, vp, P(hl_cluster_params), vp, u32, vp, vp, vp, u32, vp]),
    "hl_query_overflow": (None, [vp, P(u32)]),
    "hl_gbuffer_pack": (None, [vp, P(hl_surface), P(hl_gbuffer)]),
    "hl_deferred_forward": (None, [vp, vp, P(DeferredGlobals), P(hl_gbuffer), vp, u32,
                                   P(hl_light_grid), vp, u32, u32]),
    "hl_deferred_forward_noalloc": (None, [vp, vp, P(DeferredGlobals), P(hl_gbuffer), vp, u32,
                                           P(hl_light_grid), vp, u32, u32]),
    "hl_deferred_forward_shadowed": (None, [vp, vp, P(DeferredGlobals), P(hl_gbuffer), vp, u32,
                                            P(hl_light_grid), P(hl_shadow_maps), vp, u32, u32]),
    "hl_deferred_forward_ex": (None, [vp, vp, P(DeferredGlobals), P(hl_gbuffer), vp, u32, P(hl_light_grid),
                                      P(hl_shadow_maps), P(hl_texture), vp, u32, u32]),
    "hl_deferred_reflection": (None, [vp, vp, P(DeferredGlobals), P(hl_gbuffer), P(hl_reflection_inputs), vp,
                                       u32, u32]),
    "hl_count_pixel_light_pairs": (None, [vp, vp, P(DeferredGlobals), P(hl_gbuffer), vp, u32,
                                          P(hl_light_grid), P(C.c_uint64)]),
    "hl_shade_walk_stats": (None, [vp, vp, P(DeferredGlobals), P(hl_gbuffer), vp, u32, P(hl_light_grid),
                                   P(C.c_uint64), P(C.c_uint64)]),
    "hl_forward_lit": (None, [vp, vp, P(ForwardLitGlobals), vp, u32, vp, u32, vp, vp, u32,
                              P(hl_light_grid), vp, u32]),
    "hl_forward_lit_count_pairs": (None, [vp, vp, P(ForwardLitGlobals), vp, u32, vp, u32, vp, vp, u32,
                                          P(hl_light_grid), P(C.c_uint64)]),
    "hl_deferred_backward": (None, [vp, vp, P(DeferredGlobals), P(hl_gbuffer), vp, u32,
                                    P(hl_light_grid), vp, u32, vp, vp, vp]),
    "hl_deferred_forward_backward_l2": (None, [vp, vp, P(DeferredGlobals), P(hl_gbuffer), vp, u32,
                                               P(hl_light_grid), vp, u32, vp, u32, vp, vp, vp, vp]),
    "hl_batch_plan_create": (None, [vp, vp, u32, P(hl_view_desc), u32, P(hl_autotune_opts), P(vp)]),
    "hl_batch_capture": (None, [vp]),
    "hl_batch_launch": (None, [vp]),
    "hl_batch_get_config": (None, [vp, u32, P(hl_view_config)]),
    "hl_plan_destroy": (None, [vp]),
    "hl_external_import": (None, [vp, C.c_int, C.c_uint64, u32, P(vp)]),
    "hl_external_map": (None, [vp, C.c_uint64, C.c_uint64, P(vp)]),
    "hl_external_release": (None, [vp]),
}

# Entry points that take the context first (for error messages).
_CTX_FIRST = {n for n, (_, a) in SIGNATURES.items() if a and a[0] is vp and n.startswith("hl_")
              and n not in ("hl_context_destroy", "hl_get_stream", "hl_last_error",
                            "hl_event_destroy", "hl_batch_capture", "hl_batch_launch",
                            "hl_batch_get_config", "hl_plan_destroy", "hl_external_map",
                            "hl_external_release")}


def library_path() -> Path:
    return PKG / "libhelio_hip.so"


class Lib:
    """libhelio_hip.so with typed, status-checked entry points."""

    def __init__(self, path: Path | None = None):
        path = Path(path) if path else library_path()
        if not path.exists():
            raise FileNotFoundError(
                f"{path} is missing: build it with `python -m helio_amd.build` (hipcc, gfx950)")
        self.path = path
        self.dll = C.CDLL(str(path))
        self.raw = {}
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(self.dll, name):
                continue
            fn = getattr(self.dll, name)
            fn.argtypes = args
            fn.restype = res if res is not None else C.c_int32
            self.raw[name] = fn

    def exported(self) -> list[str]:
        return [n for n in SIGNATURES if hasattr(self.dll, n)]

    def call(self, name: str, *args, ctx=None):
        st = self.raw[name](*args)
        if SIGNATURES[name][0] is None and st != HL_OK:
            msg = ""
            if ctx is None and name in _CTX_FIRST and args:
                ctx = args[0]
            if ctx:
                m = self.raw["hl_last_error"](ctx)
                msg = m.decode() if m else ""
            if not msg:
                msg = self.raw["hl_status_string"](st).decode()
            raise HelioError(st, f"{name}: {msg}")
        return st
