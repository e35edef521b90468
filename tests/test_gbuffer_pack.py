"""The G-buffer producer's format contract (hl_gbuffer_pack against the oracle
oracle/gbuffer_pack.c, which restates gbuffer.wgsl:10-15, 324-328, 480-493).

CPU (pinning the oracle):
  * float -> half: IEEE round-to-nearest-even known answers (ties, subnormals,
    overflow), and numpy's own float16 conversion over a sweep of 10^6 values
    through the oracle's packing of an fp16 plane;
  * unorm8: known answers at the boundaries and every code's round trip
    through the exact v/255 decode;
  * F0: metallic 0 gives 0.04 exactly, metallic 1 the albedo clamped to 0.999;
    the F0 channels and the lightmap sentinel land where gbuffer.wgsl puts them.
GPU:
  * byte-for-byte equality with the oracle for the three format sets, with and
    without a specular-workflow F0, including out-of-range and non-finite
    values (NaN excluded from fp16 channels: its payload is unspecified);
  * a packed G-buffer shaded by the library equals the oracle's shading of the
    oracle-packed one within the shading tolerance.
"""
import ctypes as C

import numpy as np
import pytest

from helio_amd import abi, synth
from helio_amd.device import pack_gbuffer, surface_desc
import oracle_lib as O

FORMAT_SETS = {"ref": synth.GBUF_REF, "f16": synth.GBUF_F16, "f32": synth.GBUF_F32}


def half_bits(v):
    return int(O.oracle().or_float_to_half(C.c_float(v)))


def test_half_known_answers():
    cases = [  # a list: 0.0 and -0.0 would be one dict key
        (0.0, 0x0000), (-0.0, 0x8000), (1.0, 0x3C00), (-2.0, 0xC000), (0.1, 0x2E66), (65504.0, 0x7BFF),
        (65519.996, 0x7BFF),            # just below the rounding boundary to infinity
        (65520.0, 0x7C00),              # halfway between 65504 and 2^16: ties to even -> inf
        (1e9, 0x7C00), (float("inf"), 0x7C00), (float("-inf"), 0xFC00),
        (2.0 ** -24, 0x0001),           # smallest subnormal
        (2.0 ** -25, 0x0000),           # half of it: tie -> even (0)
        (3 * 2.0 ** -25, 0x0002),       # 1.5 ulp: tie -> even (2)
        (2.0 ** -14, 0x0400),           # smallest normal
        (2.0 ** -14 - 2.0 ** -25, 0x0400),  # rounds up into the normal range
        (1.0 + 2.0 ** -11, 0x3C00),     # tie between 0x3C00 and 0x3C01 -> even
        (1.0 + 3 * 2.0 ** -11, 0x3C02),  # tie between 0x3C01 and 0x3C02 -> even
        (1e-10, 0x0000),
    ]
    for v, want in cases:
        assert half_bits(v) == want, (v, hex(half_bits(v)), hex(want))
    assert half_bits(float("nan")) & 0x7C00 == 0x7C00 and half_bits(float("nan")) & 0x3FF


def host_surface(W, H, seed, wild=False):
    rng = np.random.default_rng(seed)
    if wild:  # out-of-range, large, tiny and signed values
        def plane():
            mags = 10.0 ** rng.uniform(-9, 6, size=(H, W, 4))
            return (rng.choice([-1.0, 1.0], size=(H, W, 4)) * mags).astype(np.float32)
        planes = [plane() for _ in range(4)]
        planes[0][..., :3] = rng.uniform(-0.3, 1.3, size=(H, W, 3))
        planes[2][..., :3] = rng.uniform(-0.3, 1.3, size=(H, W, 3))
        planes[0][0, 0, :3] = [np.inf, -np.inf, np.nan]   # unorm8 targets only
        planes[2][0, 1, :3] = [np.nan, 0.5, 1.0]
    else:
        n = rng.normal(size=(H, W, 3))
        n /= np.linalg.norm(n, axis=2, keepdims=True)
        n[..., 2] = np.abs(n[..., 2])
        planes = [np.zeros((H, W, 4), np.float32) for _ in range(4)]
        planes[0][..., :3] = rng.uniform(0.05, 0.95, size=(H, W, 3))
        planes[0][..., 3] = 1.0
        planes[1][..., :3] = n
        planes[2][..., 0] = 1.0
        planes[2][..., 1] = rng.uniform(0.05, 1.0, size=(H, W))
        planes[2][..., 2] = rng.choice([0.0, 0.3, 1.0], size=(H, W))
        planes[3][..., :3] = rng.uniform(0.0, 0.05, size=(H, W, 3))
    return [np.ascontiguousarray(p, dtype=np.float32) for p in planes]


def host_targets(W, H, gset, lightmap=True):
    fa, ta, fn, tn = synth._FMT[gset]
    arrs = {"albedo": np.zeros(H * W * ta, np.uint8), "orm": np.zeros(H * W * ta, np.uint8),
            "normal": np.zeros(H * W * tn, np.uint8), "emissive": np.zeros(H * W * tn, np.uint8),
            "lightmap": np.zeros(H * W * 4, np.uint8)}
    g = abi.hl_gbuffer()
    g.width, g.height = W, H
    for name, fmt, tb in (("albedo", fa, ta), ("orm", fa, ta), ("normal", fn, tn), ("emissive", fn, tn)):
        t = getattr(g, name)
        t.data_d, t.pitch_bytes, t.format = arrs[name].ctypes.data, W * tb, fmt
    if lightmap:
        g.lightmap_uv.data_d, g.lightmap_uv.pitch_bytes = arrs["lightmap"].ctypes.data, W * 4
        g.lightmap_uv.format = abi.HL_FMT_RG16_FLOAT
    return g, arrs


def oracle_pack(W, H, planes, gset, f0=None, lightmap=True):
    g, arrs = host_targets(W, H, gset, lightmap)
    s = surface_desc(W, H, *planes, specular_f0=f0)
    O.oracle().or_gbuffer_pack(C.byref(s), C.byref(g))
    return arrs


def test_half_matches_numpy_over_a_sweep():
    W, H = 1000, 250  # 10^6 values through the three channels of the normal plane and more
    rng = np.random.default_rng(11)
    bits = rng.integers(0x2F000000, 0x48000000, size=(H, W, 4), dtype=np.uint32)  # 2^-33 .. 2^17
    vals = bits.view(np.float32) * rng.choice([-1.0, 1.0], size=(H, W, 4)).astype(np.float32)
    planes = [np.full((H, W, 4), 0.5, np.float32), np.ascontiguousarray(vals),
              np.full((H, W, 4), 0.5, np.float32), np.ascontiguousarray(-vals)]
    arrs = oracle_pack(W, H, planes, synth.GBUF_F16)
    got = arrs["normal"].view(np.uint16).reshape(H, W, 4)[..., :3]
    with np.errstate(over="ignore"):  # values beyond 65504 become inf, as they should
        want = vals[..., :3].astype(np.float16).view(np.uint16)
        want_e = (-vals[..., :3]).astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    got_e = arrs["emissive"].view(np.uint16).reshape(H, W, 4)[..., :3]
    assert np.array_equal(got_e, want_e)


def test_unorm8_known_answers_and_round_trip():
    enc = lambda v: int(O.oracle().or_unorm8_encode(C.c_float(v)))  # noqa: E731
    assert [enc(v) for v in (0.0, -0.0, -1.0, 1.0, 2.0, float("inf"), float("-inf"), float("nan"))] == \
        [0, 0, 0, 255, 255, 255, 0, 0]
    assert enc(0.5) == 128      # 127.5 exactly: ties to even
    assert enc(0.25) == 64      # 63.75
    assert enc(np.nextafter(np.float32(0.5), np.float32(0))) == 127
    # every code survives decode (the exact v / 255 the shading kernels use) then encode
    for u in range(256):
        assert enc(float(np.float32(u) / np.float32(255))) == u


def test_f0_closed_forms_and_channel_placement():
    f0 = lambda a, m: float(O.oracle().or_f0_metallic(C.c_float(a), C.c_float(m)))  # noqa: E731
    for a in (0.0, 0.3, 0.97, 1.5):
        assert f0(a, 0.0) == float(np.float32(0.04))
        assert f0(a, 1.0) == min(float(np.float32(a)), float(np.float32(0.999)))
    assert f0(-0.5, 1.0) == 0.0
    W, H = 8, 4
    planes = host_surface(W, H, 3)
    arrs = oracle_pack(W, H, planes, synth.GBUF_F32)
    alb, orm = planes[0], planes[2]
    normal = arrs["normal"].view(np.float32).reshape(H, W, 4)
    ormt = arrs["orm"].view(np.float32).reshape(H, W, 4)
    em = arrs["emissive"].view(np.float32).reshape(H, W, 4)
    for y in range(H):
        for x in range(W):
            want = [f0(alb[y, x, k], orm[y, x, 2]) for k in range(3)]
            assert [normal[y, x, 3], ormt[y, x, 3], em[y, x, 3]] == want
    assert np.array_equal(normal[..., :3], planes[1][..., :3])
    assert np.all(arrs["lightmap"].view(np.uint16) == 0xBC00)  # (-1, -1) in fp16


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(FORMAT_SETS))
@pytest.mark.parametrize("specular", [False, True])
def test_pack_matches_oracle(ctx, name, specular):
    W, H = 333, 77  # ragged against the 64 x 4 blocks
    gset = FORMAT_SETS[name]
    planes = host_surface(W, H, 5, wild=True)
    if gset != synth.GBUF_F32:  # NaN payloads through fp16 are unspecified: keep fp16 channels finite
        for p in (planes[1], planes[3]):
            p[~np.isfinite(p)] = 1.0
    f0 = np.ascontiguousarray(np.random.default_rng(6).uniform(-0.1, 1.1, size=(H, W, 4)), np.float32) \
        if specular else None
    print(f"[gbuffer pack] {name} specular={specular}", flush=True)
    want = oracle_pack(W, H, planes, gset, f0)
    dev = [ctx.upload(p) for p in planes]
    dev_f0 = ctx.upload(f0) if specular else None
    g, host = host_targets(W, H, gset)
    outs = {}
    for name_, t in (("albedo", g.albedo), ("orm", g.orm), ("normal", g.normal), ("emissive", g.emissive),
                     ("lightmap", g.lightmap_uv)):
        outs[name_] = ctx.alloc(host[name_].size)
        outs[name_].fill_u32(0xDEADBEEF)
        t.data_d = outs[name_].ptr
    pack_gbuffer(ctx, surface_desc(W, H, *dev, specular_f0=dev_f0), g)
    ctx.synchronize()
    for k, arr in outs.items():
        got = arr.download(np.uint8)[: want[k].size]
        assert np.array_equal(got, want[k]), k


@pytest.mark.gpu
def test_pack_then_shade(ctx):
    """A G-buffer written by hl_gbuffer_pack feeds the shading kernels: the
    library's radiance from it equals the oracle's from the oracle-packed one."""
    from test_gpu_parity import check_radiance
    from helio_amd.device import DeferredLightPass, LightCullPass, upload_scene
    sc = synth.make_scene("c1_720p_64")  # reference formats, 64 lights
    W, H = sc.width, sc.height
    planes = host_surface(W, H, 9)
    packed = oracle_pack(W, H, planes, synth.GBUF_REF, lightmap=False)
    for k in ("albedo", "orm", "normal", "emissive"):
        cur = getattr(sc, k)
        setattr(sc, k, np.ascontiguousarray(packed[k].view(cur.dtype).reshape(cur.shape)))
    ds = upload_scene(ctx, sc)
    for k in ("albedo", "orm", "normal", "emissive"):
        ds.planes[k].fill_u32(0)
    dev = [ctx.upload(p) for p in planes]
    pack_gbuffer(ctx, surface_desc(W, H, *dev), ds.gbuffer)
    cull = LightCullPass(ctx, W, H)
    shade = DeferredLightPass(ctx, W, H)
    cull.prepare(ds)
    shade.prepare(ds)
    cull.execute(ds)
    shade.execute(ds, cull.publish())
    out = shade.output()
    lists, counts = O.cull_tiles(sc)
    ref, _ = O.shade(sc, O.tile_grid_host(sc, lists.ravel(), counts))
    check_radiance(out, ref)


def test_pack_rejects_bad_arguments_without_a_gpu():
    """Argument checks need no device: a null context is refused at once."""
    from helio_amd.device import lib
    s = abi.hl_surface()
    g = abi.hl_gbuffer()
    assert lib().raw["hl_gbuffer_pack"](None, C.byref(s), C.byref(g)) == abi.HL_EINVAL
