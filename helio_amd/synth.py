"""Synthetic workloads of the deferred-lighting path (host side).

Wraps libhelio_synth.so (helio_amd/csrc/synth.c) and names the five workloads
the implementation is measured and checked on:

  c1_720p_64    1280x720, reference G-buffer formats, 64 point lights,
                tiles (16 px) and 16x9x24 log clusters
  c2_1080p_1k   1920x1080, reference formats, 1024 lights (25 % spot),
                16x9x24 clusters capped at 256
  c3_4k_16k     3840x2160, fp16 G-buffer + fp32 depth, 16384 point lights,
                32x18x48 clusters, unbounded
  c4_512_adj    512x512, fp32 G-buffer, 256 point lights (adjoint workload)
  c5_multiview  48 cube-probe faces + 2 XR eyes over an 8192-light pool

Scene layout, seeds and camera poses are fixed here (the workload
descriptions leave them open); every array is a pure function of the seed.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace
from pathlib import Path

import numpy as np

from . import abi

PKG = Path(__file__).resolve().parent
SEED_BASE = 0x48454C494F000000


class hl_synth_config(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("gbuf_format", C.c_uint32),
        ("num_lights", C.c_uint32), ("spot_fraction", C.c_float),
        ("spot_outer_min_deg", C.c_float), ("spot_outer_max_deg", C.c_float),
        ("spot_inner_ratio", C.c_float), ("light_box_center", C.c_float * 3),
        ("light_box_size", C.c_float * 3), ("range_min", C.c_float), ("range_max", C.c_float),
        ("intensity_min", C.c_float), ("intensity_max", C.c_float), ("color_min", C.c_float),
        ("color_max", C.c_float), ("eye", C.c_float * 3), ("target", C.c_float * 3),
        ("fov_y_deg", C.c_float), ("near_plane", C.c_float), ("far_plane", C.c_float),
        ("num_boxes", C.c_uint32), ("box_field_half", C.c_float), ("box_size_min", C.c_float),
        ("box_size_max", C.c_float), ("box_height_min", C.c_float), ("box_height_max", C.c_float),
        ("depth_noise", C.c_float), ("_pad", C.c_uint32), ("seed", C.c_uint64),
    ]


GBUF_REF, GBUF_F16, GBUF_F32 = 0, 1, 2
_FMT = {  # gbuf_format -> (albedo/orm format, texel bytes, normal/emissive format, texel bytes)
    GBUF_REF: (abi.HL_FMT_RGBA8_UNORM, 4, abi.HL_FMT_RGBA16_FLOAT, 8),
    GBUF_F16: (abi.HL_FMT_RGBA16_FLOAT, 8, abi.HL_FMT_RGBA16_FLOAT, 8),
    GBUF_F32: (abi.HL_FMT_RGBA32_FLOAT, 16, abi.HL_FMT_RGBA32_FLOAT, 16),
}


@dataclass(frozen=True)
class SceneSpec:
    name: str
    width: int
    height: int
    gbuf_format: int
    num_lights: int
    light_box_center: tuple = (0.0, 5.0, 0.0)
    light_box_size: tuple = (40.0, 10.0, 40.0)
    range_min: float = 2.0
    range_max: float = 10.0
    intensity_min: float = 0.5
    intensity_max: float = 5.0
    color_min: float = 0.5
    color_max: float = 1.0
    spot_fraction: float = 0.0
    spot_outer_min_deg: float = 10.0
    spot_outer_max_deg: float = 45.0
    spot_inner_ratio: float = 0.8
    eye: tuple = (0.0, 8.0, 32.0)
    target: tuple = (0.0, 0.0, 0.0)
    fov_y_deg: float = 60.0
    near: float = 0.1
    far: float = 1000.0
    num_boxes: int = 40
    box_field_half: float = 20.0
    box_size_min: float = 1.0
    box_size_max: float = 4.0
    box_height_min: float = 0.5
    box_height_max: float = 6.0
    depth_noise: float = 0.05
    seed: int = SEED_BASE
    # light grid of the workload
    clusters: tuple = (16, 9, 24)
    cluster_cap: int = 0
    ambient_color: tuple = (0.5, 0.5, 0.6)   # DeferredLightPass::prepare fallback (lib.rs:842)
    ambient_intensity: float = 1.0

    def config(self) -> hl_synth_config:
        c = hl_synth_config()
        c.width, c.height, c.gbuf_format, c.num_lights = self.width, self.height, self.gbuf_format, self.num_lights
        c.spot_fraction = self.spot_fraction
        c.spot_outer_min_deg, c.spot_outer_max_deg = self.spot_outer_min_deg, self.spot_outer_max_deg
        c.spot_inner_ratio = self.spot_inner_ratio
        c.light_box_center[:] = self.light_box_center
        c.light_box_size[:] = self.light_box_size
        c.range_min, c.range_max = self.range_min, self.range_max
        c.intensity_min, c.intensity_max = self.intensity_min, self.intensity_max
        c.color_min, c.color_max = self.color_min, self.color_max
        c.eye[:] = self.eye
        c.target[:] = self.target
        c.fov_y_deg, c.near_plane, c.far_plane = self.fov_y_deg, self.near, self.far
        c.num_boxes, c.box_field_half = self.num_boxes, self.box_field_half
        c.box_size_min, c.box_size_max = self.box_size_min, self.box_size_max
        c.box_height_min, c.box_height_max = self.box_height_min, self.box_height_max
        c.depth_noise = self.depth_noise
        c.seed = self.seed
        return c


SPECS: dict[str, SceneSpec] = {
    "c1_720p_64": SceneSpec("c1_720p_64", 1280, 720, GBUF_REF, 64, seed=SEED_BASE + 16 * 1),
    "c2_1080p_1k": SceneSpec("c2_1080p_1k", 1920, 1080, GBUF_REF, 1024, spot_fraction=0.25,
                             light_box_center=(0.0, 7.5, 0.0), light_box_size=(80.0, 15.0, 80.0),
                             eye=(0.0, 10.0, 55.0), box_field_half=40.0, num_boxes=60,
                             cluster_cap=256, seed=SEED_BASE + 16 * 2),
    "c3_4k_16k": SceneSpec("c3_4k_16k", 3840, 2160, GBUF_F16, 16384,
                           light_box_center=(0.0, 15.0, 0.0), light_box_size=(200.0, 30.0, 200.0),
                           range_min=1.0, range_max=8.0, eye=(0.0, 12.0, 115.0),
                           num_boxes=96, box_field_half=100.0, box_size_min=2.0, box_size_max=10.0,
                           box_height_min=1.0, box_height_max=15.0, clusters=(32, 18, 48),
                           seed=SEED_BASE + 16 * 3),
    "c4_512_adj": SceneSpec("c4_512_adj", 512, 512, GBUF_F32, 256, seed=SEED_BASE + 16 * 4),
}
# The 4K/16k lights and grid seen from 420 m out with a 90-degree field of
# view: far froxels span the light box, so cluster lists run from 0 to 2000+
# indices (the long-list regime of the 16384-light configuration).
SPECS["c3_4k_16k_far"] = replace(SPECS["c3_4k_16k"], name="c3_4k_16k_far", eye=(0.0, 60.0, 420.0),
                                 fov_y_deg=90.0)


def spec(name: str, **over) -> SceneSpec:
    return replace(SPECS[name], **over) if over else SPECS[name]


_LIB = None


def synth_lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        p = PKG / "libhelio_synth.so"
        if not p.exists():
            raise FileNotFoundError(f"{p} missing: run `python -m helio_amd.build`")
        lib = C.CDLL(str(p))
        lib.hl_synth_camera.argtypes = [C.POINTER(hl_synth_config), C.POINTER(abi.GpuCameraUniforms)]
        lib.hl_synth_lights.argtypes = [C.POINTER(hl_synth_config), C.c_void_p]
        lib.hl_synth_gbuffer.argtypes = [C.POINTER(hl_synth_config)] + [C.c_void_p] * 5
        lib.hl_synth_subset.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
        lib.hl_synth_shadow_maps.argtypes = [C.POINTER(hl_synth_config), C.c_void_p, C.c_uint32,
                                             C.POINTER(C.c_float), C.c_uint32, C.c_uint32] + [C.c_void_p] * 3
        lib.hl_synth_abi_size.restype = C.c_uint32
        assert lib.hl_synth_abi_size() == C.sizeof(hl_synth_config)
        _LIB = lib
    return _LIB


@dataclass
class HostScene:
    """Host-memory inputs of one view: camera, lights, globals and G-buffer planes."""
    spec: SceneSpec
    camera: abi.GpuCameraUniforms
    lights: np.ndarray              # GPU_LIGHT_DTYPE
    globals: abi.DeferredGlobals
    albedo: np.ndarray
    normal: np.ndarray
    orm: np.ndarray
    emissive: np.ndarray
    depth: np.ndarray               # float32 [H, W]
    formats: tuple = field(default=(0, 0, 0, 0))
    texel: tuple = field(default=(0, 0))

    @property
    def width(self) -> int:
        return self.spec.width

    @property
    def height(self) -> int:
        return self.spec.height

    def gbuffer_host(self) -> abi.hl_gbuffer:
        """hl_gbuffer whose pointers are the host arrays (for the CPU oracle)."""
        g = abi.hl_gbuffer()
        g.width, g.height = self.width, self.height
        fa, ta, fn, tn = self.formats
        for name, arr, fmt, tb in (("albedo", self.albedo, fa, ta), ("orm", self.orm, fa, ta),
                                   ("normal", self.normal, fn, tn), ("emissive", self.emissive, fn, tn)):
            t = getattr(g, name)
            t.data_d, t.pitch_bytes, t.format = arr.ctypes.data, self.width * tb, fmt
        g.depth.data_d, g.depth.pitch_bytes, g.depth.format = self.depth.ctypes.data, self.width * 4, abi.HL_FMT_R32_FLOAT
        return g


def make_globals(s: SceneSpec) -> abi.DeferredGlobals:
    g = abi.DeferredGlobals()
    g.light_count = s.num_lights
    g.ambient_intensity = s.ambient_intensity
    g.ambient_color[:] = (*s.ambient_color, 1.0)
    g.csm_splits[:] = (16.0, 80.0, 300.0, 1400.0)
    g.num_tiles_x = (s.width + 15) // 16
    g.enable_reflections = 1
    return g


def make_camera(s: SceneSpec) -> abi.GpuCameraUniforms:
    cam = abi.GpuCameraUniforms()
    cfg = s.config()
    if synth_lib().hl_synth_camera(C.byref(cfg), C.byref(cam)) != 0:
        raise ValueError("bad scene spec")
    return cam


def make_lights(s: SceneSpec) -> np.ndarray:
    lights = np.zeros(s.num_lights, dtype=abi.GPU_LIGHT_DTYPE)
    cfg = s.config()
    if synth_lib().hl_synth_lights(C.byref(cfg), lights.ctypes.data if s.num_lights else None) != 0:
        raise ValueError("bad light spec")
    return lights


def make_scene(s: SceneSpec | str, **over) -> HostScene:
    if isinstance(s, str):
        s = spec(s, **over)
    elif over:
        s = replace(s, **over)
    fa, ta, fn, tn = _FMT[s.gbuf_format]
    W, H = s.width, s.height
    alb = np.zeros((H, W, ta), dtype=np.uint8)
    orm = np.zeros((H, W, ta), dtype=np.uint8)
    nrm = np.zeros((H, W, tn), dtype=np.uint8)
    emi = np.zeros((H, W, tn), dtype=np.uint8)
    dep = np.zeros((H, W), dtype=np.float32)
    cfg = s.config()
    r = synth_lib().hl_synth_gbuffer(C.byref(cfg), alb.ctypes.data, nrm.ctypes.data, orm.ctypes.data,
                                     emi.ctypes.data, dep.ctypes.data)
    if r != 0:
        raise ValueError("bad G-buffer spec")
    return HostScene(spec=s, camera=make_camera(s), lights=make_lights(s), globals=make_globals(s),
                     albedo=alb, normal=nrm, orm=orm, emissive=emi, depth=dep,
                     formats=(fa, ta, fn, tn), texel=(ta, tn))


def perturb_for_adjoint(gt: HostScene, seed: int = 7) -> HostScene:
    """Initial guess of the differentiable-shading workload from ground truth
    (fp32 G-buffer): albedo + N(0,0.1) clamped to [0,1], roughness + N(0,0.1)
    clamped to [0.05,1], light positions + N(0,0.5), intensities x U[0.7,1.3].
    F0 is re-resolved from the perturbed albedo and metallic as the G-buffer
    producer does (gbuffer.wgsl:324-328) and packed into normal.a / orm.a /
    emissive.a."""
    if gt.spec.gbuf_format != GBUF_F32:
        raise ValueError("the adjoint workload uses an fp32 G-buffer")
    W, H = gt.width, gt.height
    rng = np.random.Generator(np.random.PCG64(seed))
    g = make_scene(gt.spec)
    alb = g.albedo.view(np.float32).reshape(H, W, 4)
    orm = g.orm.view(np.float32).reshape(H, W, 4)
    alb[..., :3] = np.clip(alb[..., :3] + rng.normal(0, 0.1, (H, W, 3)), 0.0, 1.0).astype(np.float32)
    orm[..., 1] = np.clip(orm[..., 1] + rng.normal(0, 0.1, (H, W)), 0.05, 1.0).astype(np.float32)
    m = orm[..., 2:3]
    f0 = np.clip(np.float32(0.04) * (np.float32(1) - m) + alb[..., :3] * m, 0.0, np.float32(0.999)).astype(np.float32)
    nrm = g.normal.view(np.float32).reshape(H, W, 4)
    emi = g.emissive.view(np.float32).reshape(H, W, 4)
    nrm[..., 3], orm[..., 3], emi[..., 3] = f0[..., 0], f0[..., 1], f0[..., 2]
    lights = g.lights.copy()
    n = len(lights)
    lights["position_range"][:, :3] += rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    lights["color_intensity"][:, 3] *= rng.uniform(0.7, 1.3, n).astype(np.float32)
    g.lights = lights
    return g


def light_subset(seed: int, view: int, pool: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    if synth_lib().hl_synth_subset(seed, view, pool, n, out.ctypes.data if n else None) != 0:
        raise ValueError("bad subset")
    return out


def cluster_params(lib: abi.Lib, scene: HostScene, grid=None, cap=None) -> abi.hl_cluster_params:
    gx, gy, gz = grid or scene.spec.clusters
    p = abi.hl_cluster_params()
    lib.call("hl_cluster_params_init", C.byref(scene.camera), scene.width, scene.height, gx, gy, gz,
             scene.spec.cluster_cap if cap is None else cap, C.byref(p))
    return p


# ----------------------------------------------------------- multi-view set
def multiview_specs(seed: int = SEED_BASE + 16 * 5, small: bool = False) -> tuple[list[SceneSpec], list[np.ndarray], SceneSpec]:
    """48 probe faces (16 each at 128, 256, 512 px) + 2 XR eyes at 2064x2208 over
    an 8192-light pool; each view takes a log-uniform [32, 4096] subset.
    small=True shrinks every view 8x per side (tests)."""
    pool = SceneSpec("c5_pool", 1, 1, GBUF_REF, 8192, light_box_center=(0.0, 10.0, 0.0),
                     light_box_size=(100.0, 20.0, 100.0), range_min=1.0, range_max=8.0, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    faces = [(0.0, 0.0, -1.0), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, -1.0, 0.0)]
    views: list[SceneSpec] = []
    subsets: list[np.ndarray] = []
    sizes = [128] * 16 + [256] * 16 + [512] * 16
    probe = None
    for i, sz in enumerate(sizes):
        if i % 6 == 0:
            probe = (float(rng.uniform(-40, 40)), float(rng.uniform(2, 15)), float(rng.uniform(-40, 40)))
        f = faces[i % 6]
        tgt = (probe[0] + f[0], probe[1] + f[1], probe[2] + f[2])
        s = max(8, sz // 8) if small else sz
        views.append(SceneSpec(f"face{i}", s, s, GBUF_REF, 0, eye=probe, target=tgt, fov_y_deg=90.0,
                               num_boxes=60, box_field_half=50.0, seed=seed + 1000 + i))
    head = (0.0, 1.7, 30.0)
    for e, dx in enumerate((-0.032, 0.032)):
        w, h = (2064, 2208) if not small else (2064 // 8, 2208 // 8)
        views.append(SceneSpec(f"eye{e}", w, h, GBUF_REF, 0, eye=(head[0] + dx, head[1], head[2]),
                               target=(head[0] + dx, 0.0, 0.0), fov_y_deg=100.0, num_boxes=60,
                               box_field_half=50.0, seed=seed + 2000 + e))
    for i in range(len(views)):
        n = int(round(math.exp(rng.uniform(math.log(32), math.log(4096)))))
        subsets.append(light_subset(seed, i, pool.num_lights, n))
    return views, subsets, pool


# ------------------------------------------------------------ forward-lit
def _decode_plane(arr: np.ndarray, fmt: int, H: int, W: int) -> np.ndarray:
    if fmt == abi.HL_FMT_RGBA8_UNORM:
        return arr.reshape(H, W, 4).astype(np.float32) / np.float32(255.0)
    if fmt == abi.HL_FMT_RGBA16_FLOAT:
        return arr.view(np.float16).reshape(H, W, 4).astype(np.float32)
    return arr.view(np.float32).reshape(H, W, 4)


def make_fragments(scene: HostScene, order: str = "tiles", seed: int = 0) -> np.ndarray:
    """One forward-shaded fragment (abi.FRAGMENT_DTYPE) per non-sky pixel of
    the scene: world position at the pixel centre from its depth (computed in
    double, so it projects into its pixel), the unit normal, albedo and alpha,
    ao, roughness clamped to [0.045, 1] and metallic of the G-buffer, F0 of the
    metallic workflow and the emissive colour -- what a rasteriser plus the
    material evaluation would hand the forward-lit shader for that surface.
    order: "tiles" (8x8-pixel blocks in raster order, as a tiled rasteriser
    emits them), "rows" (raster order) or "shuffled" (a random permutation)."""
    H, W = scene.height, scene.width
    fa, _, fn, _ = scene.formats
    ys, xs = np.nonzero(scene.depth < 1.0)
    if order == "tiles":
        key = ((ys // 8) * ((W + 7) // 8) + xs // 8) * 64 + (ys % 8) * 8 + xs % 8
        o = np.argsort(key, kind="stable")
        ys, xs = ys[o], xs[o]
    elif order == "shuffled":
        o = np.random.default_rng(seed).permutation(len(ys))
        ys, xs = ys[o], xs[o]
    elif order != "rows":
        raise ValueError(f"unknown fragment order {order!r}")
    n = len(ys)
    f = np.zeros(n, dtype=abi.FRAGMENT_DTYPE)
    ivp = np.array(scene.camera.inv_view_proj, dtype=np.float64).reshape(4, 4).T  # column-major
    d = scene.depth[ys, xs].astype(np.float64)
    ndc = np.stack([((xs + 0.5) / W) * 2.0 - 1.0, 1.0 - ((ys + 0.5) / H) * 2.0, d, np.ones(n)])
    h = ivp @ ndc
    f["world_pos"] = (h[:3] / h[3]).T.astype(np.float32)
    f["pixel"] = xs.astype(np.uint32) | (ys.astype(np.uint32) << 16)
    alb = _decode_plane(scene.albedo, fa, H, W)[ys, xs]
    nrm = _decode_plane(scene.normal, fn, H, W)[ys, xs].astype(np.float64)
    orm = _decode_plane(scene.orm, fa, H, W)[ys, xs]
    emi = _decode_plane(scene.emissive, fn, H, W)[ys, xs]
    nn = nrm[:, :3] / np.linalg.norm(nrm[:, :3], axis=1, keepdims=True)
    f["normal"] = nn.astype(np.float32)
    f["alpha"] = alb[:, 3]
    f["albedo"] = alb[:, :3]
    f["ao"] = orm[:, 0]
    f["roughness"] = np.clip(orm[:, 1], np.float32(0.045), np.float32(1.0))
    m = orm[:, 2:3]
    f["metallic"] = m[:, 0]
    f["specular_f0"] = np.clip(np.float32(0.04) * (np.float32(1.0) - m) + alb[:, :3] * m,
                               np.float32(0.0), np.float32(0.999))
    f["emissive"] = emi[:, :3]
    return f


def direction_from_rotation_f64(rot_deg: np.ndarray) -> np.ndarray:
    """-Z rotated by Ry(yaw) Rx(pitch) Rz(roll), rotation = [pitch, yaw, roll]
    in degrees, in double (rotation matrices, not the shader's quaternions)."""
    r = np.radians(np.asarray(rot_deg, dtype=np.float64))
    p, y, z = r[..., 0], r[..., 1], r[..., 2]
    # R = Ry Rx Rz applied to (0, 0, -1): minus the third column of R
    cx, sx, cy, sy = np.cos(p), np.sin(p), np.cos(y), np.sin(y)
    col = np.stack([sy * cx, -sx, cy * cx], axis=-1)  # Ry Rx e_z (Rz leaves e_z)
    return -col


def make_light_transforms(lights: np.ndarray, seed: int = 0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """SceneDB-style light transforms for a light array: a transform table of
    2n + 3 entries of which the lights own a random subset (entity indices),
    each with the light's position and a random rotation.  Returns (lights
    with direction_outer.xyz set to the rotated -Z, so that the clustered
    cull's cone test sees the direction the forward shader uses; entity
    indices; transforms)."""
    n = len(lights)
    rng = np.random.default_rng(seed)
    rot = rng.uniform(-180.0, 180.0, size=(n, 3)).astype(np.float32)
    ntr = 2 * n + 3
    entity = rng.permutation(ntr)[:n].astype(np.uint32)
    tr = np.zeros(ntr, dtype=abi.TRANSFORM_DTYPE)
    tr["scale"] = 1.0
    tr["position"][entity] = lights["position_range"][:, :3]
    tr["rotation"][entity] = rot
    out = lights.copy()
    out["direction_outer"][:, :3] = direction_from_rotation_f64(rot).astype(np.float32)
    return out, entity, tr


def forward_globals(scene: HostScene) -> abi.ForwardLitGlobals:
    """ForwardLitGlobals as ForwardLitPass::prepare fills them
    (crates/passes/3d/helio-pass-forward-lit/src/lib.rs:382-406)."""
    g = abi.ForwardLitGlobals()
    hg = scene.globals
    g.frame = hg.frame
    g.delta_time = hg.delta_time
    g.light_count = len(scene.lights)
    g.ambient_intensity = hg.ambient_intensity
    g.ambient_color[:] = (hg.ambient_color[0], hg.ambient_color[1], hg.ambient_color[2], 1.0)
    g.num_tiles_x = (scene.width + 15) // 16
    g.num_tiles_y = (scene.height + 15) // 16
    g.screen_width = float(scene.width)
    g.screen_height = float(scene.height)
    return g


# ----------------------------------------------------------------- shadows
# ShadowConfig::from_quality (crates/libhelio/src/shadow.rs:71-196): per cascade
# (split_distance, depth_bias, filter_radius, pcss_light_size), then enable_pcss,
# pcss_blocker_samples, pcss_filter_samples, pcf_sample_count.
_SHADOW_PRESETS = {
    "low": ([(16.0, 0.00010, 1.0, 0.0), (80.0, 0.00015, 1.5, 0.0), (300.0, 0.00020, 2.0, 0.0),
             (1400.0, 0.00030, 3.0, 0.0)], 0, 4, 4, 8),
    "medium": ([(16.0, 0.00010, 1.5, 0.0), (80.0, 0.00015, 2.0, 0.0), (300.0, 0.00020, 2.5, 0.0),
                (1400.0, 0.00030, 3.5, 0.0)], 0, 8, 8, 12),
    "high": ([(16.0, 0.00008, 2.0, 2.0), (80.0, 0.00012, 2.5, 4.0), (300.0, 0.00018, 3.0, 8.0),
              (1400.0, 0.00025, 4.0, 16.0)], 1, 8, 16, 12),
    "ultra": ([(16.0, 0.00008, 2.5, 2.0), (80.0, 0.00012, 3.0, 4.0), (300.0, 0.00018, 3.5, 8.0),
               (1400.0, 0.00025, 4.5, 16.0)], 1, 16, 32, 16),
}


def shadow_config(quality: str = "medium") -> abi.ShadowConfig:
    casc, pcss, nb, nf, npcf = _SHADOW_PRESETS[quality]
    c = abi.ShadowConfig()
    for i, (split, bias, radius, size) in enumerate(casc):
        c.cascades[i].split_distance, c.cascades[i].depth_bias = split, bias
        c.cascades[i].filter_radius, c.cascades[i].pcss_light_size = radius, size
    c.enable_pcss, c.pcss_blocker_samples, c.pcss_filter_samples, c.pcf_sample_count = pcss, nb, nf, npcf
    return c


@dataclass
class HostShadows:
    """Shadow atlases (dynamic, static: [layers, size, size] float32), one light
    matrix per layer ([layers, 16], column-major) and the ShadowConfig."""
    atlas: np.ndarray
    static_atlas: np.ndarray
    matrices: np.ndarray
    config: abi.ShadowConfig

    @property
    def size(self) -> int:
        return self.atlas.shape[1]

    @property
    def layers(self) -> int:
        return self.atlas.shape[0]

    def host(self) -> abi.hl_shadow_maps:
        """hl_shadow_maps whose pointers are the host arrays (for the CPU oracle)."""
        h = abi.hl_shadow_maps()
        h.atlas_d, h.static_atlas_d = self.atlas.ctypes.data, self.static_atlas.ctypes.data
        h.matrices_d = self.matrices.ctypes.data
        h.num_layers, h.atlas_size, h.config = self.layers, self.size, self.config
        return h


def add_shadow_casters(scene: HostScene, n_point: int = 8, n_spot: int = 8, directional: bool = True,
                       size: int = 256, quality: str = "medium", seed: int = 11) -> HostShadows:
    """Turns the first lights of the scene into shadow casters -- light 0 a
    directional sun (4 cascades), then n_point point lights (6 faces each) and
    n_spot spot lights (one layer each), all below MAX_SHADOW_LIGHTS = 42 --
    and renders their shadow maps from the scene's boxes and ground
    (libhelio_synth: one ray per texel)."""
    lights = scene.lights
    rng = np.random.default_rng(seed)
    first = 1 if directional else 0
    assert first + n_point + n_spot <= min(42, len(lights))
    layer = 0
    if directional:
        d = np.array([0.35, -1.0, 0.25]); d /= np.linalg.norm(d)
        lights[0]["light_type"] = abi.HL_LIGHT_DIRECTIONAL
        lights[0]["direction_outer"] = (*d, 0.0)
        lights[0]["color_intensity"] = (1.0, 0.95, 0.85, 1.5)
        lights[0]["shadow_index"] = 0
        layer = 4
    for i in range(first, first + n_point):
        lights[i]["light_type"] = abi.HL_LIGHT_POINT
        lights[i]["direction_outer"] = (0.0, -1.0, 0.0, 0.0)
        lights[i]["shadow_index"] = layer
        layer += 6
    for i in range(first + n_point, first + n_point + n_spot):
        d = rng.normal(size=3); d[1] = -abs(d[1]) - 0.5; d /= np.linalg.norm(d)
        outer = np.radians(rng.uniform(25.0, 45.0))
        lights[i]["light_type"] = abi.HL_LIGHT_SPOT
        lights[i]["direction_outer"] = (*d, np.cos(outer))
        lights[i]["inner_angle"] = np.cos(outer * 0.8)
        lights[i]["shadow_index"] = layer
        layer += 1
    atlas = np.zeros((layer, size, size), np.float32)
    static = np.zeros_like(atlas)
    mats = np.zeros((layer, 16), np.float32)
    csm = (C.c_float * 4)(*scene.globals.csm_splits)
    cfg = scene.spec.config()
    if synth_lib().hl_synth_shadow_maps(C.byref(cfg), lights.ctypes.data, len(lights), csm, size, layer,
                                        atlas.ctypes.data, static.ctypes.data, mats.ctypes.data) != 0:
        raise ValueError("bad shadow map request")
    return HostShadows(atlas, static, mats, shadow_config(quality))


# --------------------------------------------------------- radiance cascades
def make_rc_probes(scene: HostScene, fmt: int = abi.HL_FMT_RGBA16_FLOAT, seed: int = 13,
                   box: tuple | None = None) -> np.ndarray:
    """A synthetic rc_cascade0 probe atlas (64 x 1024 texels: 16^3 probes of
    4 x 4 directions) with smooth positive radiance, and the scene's globals
    switched to radiance-cascade GI over box ((min xyz), (max xyz); default:
    the central part of the scene, so that the volume fade shows)."""
    rng = np.random.default_rng(seed)
    px, py, pz, dx, dy = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), np.arange(4), np.arange(4),
                                     indexing="ij")
    base = rng.uniform(0.0, 0.6, size=(3, 4, 4))
    val = np.zeros((16, 16, 16, 4, 4, 3))
    for c in range(3):
        val[..., c] = base[c][dx, dy] * (0.5 + 0.5 * np.sin(0.4 * px + 0.3 * py + 0.5 * pz + c))
    atlas = np.zeros((1024, 64, 4))
    # texel (x, y) = (px * 4 + dx, (py * 16 + pz) * 4 + dy)
    atlas[((py * 16 + pz) * 4 + dy).ravel(), (px * 4 + dx).ravel(), :3] = val.reshape(-1, 3)
    atlas[..., 3] = 1.0
    g = scene.globals
    lo, hi = box if box is not None else ((-15.0, -1.0, -15.0), (15.0, 12.0, 15.0))
    g.rc_world_min[:] = (*lo, 0.0)
    g.rc_world_max[:] = (*hi, 0.0)
    g.has_rc_gi = 1
    dt = np.float16 if fmt == abi.HL_FMT_RGBA16_FLOAT else np.float32
    return np.ascontiguousarray(atlas.astype(dt))


def rc_texture_host(atlas: np.ndarray) -> abi.hl_texture:
    t = abi.hl_texture()
    t.data_d, t.pitch_bytes = atlas.ctypes.data, atlas.shape[1] * atlas.itemsize * 4
    t.format = abi.HL_FMT_RGBA16_FLOAT if atlas.dtype == np.float16 else abi.HL_FMT_RGBA32_FLOAT
    return t


# ------------------------------------------------------------- reflections
def cube_dirs(res: int) -> np.ndarray:
    """Unit directions through the texel centres of the six faces (+X -X +Y
    -Y +Z -Z) of a res x res cube, [6, res, res, 3]: the inverse of the
    sampler's face table (sc, tc on the major axis)."""
    c = (np.arange(res) + 0.5) / res * 2.0 - 1.0
    tc, sc = np.meshgrid(c, c, indexing="ij")  # row: tc, column: sc
    one = np.ones_like(sc)
    faces = [np.stack([one, -tc, -sc], -1), np.stack([-one, -tc, sc], -1),
             np.stack([sc, one, tc], -1), np.stack([sc, -one, -tc], -1),
             np.stack([sc, -tc, one], -1), np.stack([-sc, -tc, -one], -1)]
    d = np.stack(faces)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


@dataclass
class HostReflections:
    """Inputs of hl_deferred_reflection on the host: the cube array (cube 0
    the skylight, cube i + 1 capture i's), the capture records and the
    optional screen textures."""
    cubes: np.ndarray
    face_size: int
    mip_levels: int
    num_cubes: int
    format: int
    captures: np.ndarray
    ssr: np.ndarray | None = None
    planar: np.ndarray | None = None
    rc: np.ndarray | None = None

    def env_host(self) -> abi.hl_cube_array:
        e = abi.hl_cube_array()
        e.data_d, e.face_size, e.mip_levels = self.cubes.ctypes.data, self.face_size, self.mip_levels
        e.num_cubes, e.format = self.num_cubes, self.format
        return e

    def inputs_host(self) -> abi.hl_reflection_inputs:
        r = abi.hl_reflection_inputs()
        r.env = self.env_host()
        r.captures_d = self.captures.ctypes.data if len(self.captures) else None
        for name, img in (("ssr", self.ssr), ("planar", self.planar), ("rc_cascade0", self.rc)):
            if img is not None:
                t = getattr(r, name)
                t.data_d, t.pitch_bytes = img.ctypes.data, img.shape[1] * img.itemsize * 4
                t.format = abi.HL_FMT_RGBA16_FLOAT if img.dtype == np.float16 else abi.HL_FMT_RGBA32_FLOAT
        return r


def make_cube_array(num_cubes: int, face: int, mips: int, fmt: int = abi.HL_FMT_RGBA16_FLOAT,
                    seed: int = 21) -> np.ndarray:
    """A pre-filtered-looking cube array: cube c mip m holds a smooth positive
    function of the texel direction, its own tint and a per-mip fade, in the
    layout hl_cube_array documents (cubes, then mips, then faces, then rows)."""
    rng = np.random.default_rng(seed)
    parts = []
    for c in range(num_cubes):
        tint = rng.uniform(0.3, 1.5, size=3)
        k = rng.uniform(1.0, 3.0, size=3)
        ph = rng.uniform(0.0, 6.28, size=3)
        for m in range(mips):
            r = max(1, face >> m)
            d = cube_dirs(r)
            val = np.empty((6, r, r, 4))
            for ch in range(3):
                val[..., ch] = tint[ch] * (0.55 + 0.45 * np.sin(k[ch] * d[..., ch] + d[..., (ch + 1) % 3] + ph[ch])) \
                    * (1.0 - 0.06 * m)
            val[..., 3] = 1.0
            parts.append(val.reshape(-1, 4))
    arr = np.concatenate(parts)
    if fmt == abi.HL_FMT_RGBA8_UNORM:
        return np.ascontiguousarray(np.clip(np.rint(np.clip(arr, 0, 1) * 255.0), 0, 255).astype(np.uint8))
    return np.ascontiguousarray(arr.astype(np.float16 if fmt == abi.HL_FMT_RGBA16_FLOAT else np.float32))


def make_reflections(scene: HostScene, n_box: int = 4, n_sphere: int = 4, face: int = 64, mips: int = 7,
                     fmt: int = abi.HL_FMT_RGBA16_FLOAT, ssr: bool = True, planar: bool = True,
                     rc: bool = False, seed: int = 23) -> HostReflections:
    """Reflection captures around visible points of the scene (boxes with a
    random yaw, spheres; one capture without a cubemap), sorted by influence
    size smallest first (crates/libhelio/src/reflection.rs:46-49), their cube
    array, an SSR image whose confidence is zero over part of the frame, a
    planar image on up-facing pixels, and (rc) the radiance-cascade atlas.
    Switches the scene's globals to reflections on."""
    rng = np.random.default_rng(seed)
    H, W = scene.height, scene.width
    cam = scene.camera
    ivp = np.array(cam.inv_view_proj, np.float64).reshape(4, 4).T  # column-major storage
    ys, xs = np.nonzero(scene.depth < 1.0)
    pick = rng.choice(len(xs), size=n_box + n_sphere, replace=False)
    caps = np.zeros(n_box + n_sphere, abi.CAPTURE_DTYPE)
    for i, j in enumerate(pick):
        x, y = xs[j], ys[j]
        ndc = np.array([(x + 0.5) / W * 2 - 1, 1 - (y + 0.5) / H * 2, float(scene.depth[y, x]), 1.0])
        h = ivp @ ndc
        p = h[:3] / h[3]
        c = caps[i]
        c["cubemap_index"] = i + 1
        c["brightness"] = rng.uniform(0.6, 1.4)
        if i < n_box:
            yaw = rng.uniform(0, 2 * np.pi)
            rot = np.array([[np.cos(yaw), 0, np.sin(yaw)], [0, 1, 0], [-np.sin(yaw), 0, np.cos(yaw)]])
            ext = rng.uniform((3.0, 2.0, 3.0), (10.0, 6.0, 10.0))
            m = np.eye(4)
            m[:3, :3] = rot.T
            m[:3, 3] = -rot.T @ p
            c["position_radius"] = (*p, float(np.linalg.norm(ext)))
            c["extents_transition"] = (*ext, rng.uniform(0.5, 2.5))
            c["world_to_local"] = m.T.ravel()  # column-major
            c["shape"] = abi.HL_CAPTURE_SHAPE_BOX
        else:
            c["position_radius"] = (*p, rng.uniform(4.0, 14.0))
            c["world_to_local"] = np.eye(4).ravel()
            c["shape"] = abi.HL_CAPTURE_SHAPE_SPHERE
    size = np.where(caps["shape"] == abi.HL_CAPTURE_SHAPE_BOX,
                    np.linalg.norm(caps["extents_transition"][:, :3], axis=1), caps["position_radius"][:, 3])
    caps = caps[np.argsort(size, kind="stable")]
    if len(caps) > 2:
        caps[len(caps) // 2]["cubemap_index"] = -1  # not resident yet
    num_cubes = n_box + n_sphere + 1
    cubes = make_cube_array(num_cubes, face, mips, fmt, seed=seed + 1)
    g = scene.globals
    g.reflection_capture_count = len(caps)
    g.enable_reflections = 1
    g.enable_env_reflections = 1
    s_img = p_img = None
    if ssr:
        yy, xx = np.mgrid[0:H, 0:W]
        a = np.clip(np.sin(xx * 0.013) * np.cos(yy * 0.017) * 1.5, 0.0, 1.0)
        s_img = np.empty((H, W, 4), np.float32)
        s_img[..., :3] = rng.uniform(0.0, 2.0, size=(H, W, 3))
        s_img[..., 3] = a
        s_img = s_img.astype(np.float16)
    if planar:
        n = scene.normal.view(np.float16 if scene.formats[2] == abi.HL_FMT_RGBA16_FLOAT else np.float32)
        ny = n.reshape(H, W, 4)[..., 1].astype(np.float32)
        p_img = np.zeros((H, W, 4), np.float32)
        up = ny > 0.9
        p_img[up, :3] = rng.uniform(0.0, 1.5, size=(int(up.sum()), 3))
        p_img[up, 3] = 0.6
        p_img = p_img.astype(np.float16)
    rc_atlas = make_rc_probes(scene) if rc else None
    return HostReflections(cubes, face, mips, num_cubes, fmt, caps, s_img, p_img, rc_atlas)
