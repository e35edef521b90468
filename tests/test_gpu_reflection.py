"""The reflection composite on the GPU (hl_deferred_reflection, the deferred
pass's fs_reflection draw) against the FP32 restatement (oracle/reflection.c)
on the same inputs: within 1e-4 relative (+1e-6 absolute).  The kernel
follows the restatement operation for operation (IEEE division and square
root, no contraction, the same cube sampler), so the two differ only through
exp2 in env_brdf_approx; the SSR debug modes and the disabled composite are
compared bit for bit."""
import ctypes as C

import numpy as np
import pytest

from helio_amd import abi, synth
from helio_amd.device import DeviceReflections, ReflectionPass, upload_scene
import oracle_lib as O
from test_gpu_parity import check_radiance

pytestmark = pytest.mark.gpu


def run(ctx, sc, refl, base, flags=0, globals_=None):
    ds = upload_scene(ctx, sc)
    out = ctx.upload(base)
    ReflectionPass(ctx).execute(ds, DeviceReflections(ctx, refl), out, flags=flags, globals_=globals_)
    return out.download(np.float32).reshape(sc.height, sc.width, 4)


# the 4K/16k-light scene (fp16 G-buffer set) at a quarter of its pixels
C3_REDUCED = dict(width=960, height=540, num_lights=4096)


@pytest.mark.parametrize("name,fmt,rc,over", [("c1_720p_64", abi.HL_FMT_RGBA16_FLOAT, False, {}),
                                              ("c2_1080p_1k", abi.HL_FMT_RGBA8_UNORM, True, {}),
                                              ("c3_4k_16k", abi.HL_FMT_RGBA16_FLOAT, True, C3_REDUCED),
                                              ("c4_512_adj", abi.HL_FMT_RGBA32_FLOAT, True, {}),
                                              ("c4_512_adj", abi.HL_FMT_RGBA16_FLOAT, False, {})])
def test_reflection_matches_oracle(ctx, name, fmt, rc, over):
    print(f"[reflection] {name} cube format {fmt} rc={rc} {over}", flush=True)
    sc = synth.make_scene(name, **over)
    refl = synth.make_reflections(sc, n_box=5, n_sphere=5, face=64, mips=7, fmt=fmt, rc=rc)
    H, W = sc.height, sc.width
    sky = sc.depth >= 1.0
    zero = np.zeros((H, W, 4), np.float32)
    got = run(ctx, sc, refl, zero, flags=abi.HL_REFLECTION_STORE)
    ref, n = O.reflection(sc, refl, flags=abi.HL_REFLECTION_STORE)
    assert n == np.count_nonzero(~sky)
    r = check_radiance(got, ref, ~sky)
    assert np.all(got[sky] == 0.0)
    # the captures reach part of the frame (their blend differs from the skylight alone)
    g = abi.DeferredGlobals()
    C.memmove(C.byref(g), C.byref(sc.globals), C.sizeof(g))
    g.reflection_capture_count = 0
    sky_only, _ = O.reflection(sc, refl, flags=abi.HL_REFLECTION_STORE, globals_=g)
    reached = np.mean(np.any(sky_only[~sky] != ref[~sky], axis=1))
    print(f"  max rel err {r:.2e}, pixels reached by a capture {reached:.2f}")
    assert reached > 0.02
    # added over a lit frame (the pass's One + One blend; alpha kept)
    base = np.random.default_rng(3).uniform(0.0, 2.0, size=(H, W, 4)).astype(np.float32)
    got = run(ctx, sc, refl, base)
    ref, _ = O.reflection(sc, refl, out=base.copy())
    check_radiance(got, ref)
    assert np.array_equal(got[..., 3], base[..., 3])


def test_more_than_64_captures(ctx):
    print("[reflection] 70 captures (64 read)", flush=True)
    sc = synth.make_scene("c1_720p_64", width=640, height=360)
    refl = synth.make_reflections(sc, n_box=35, n_sphere=35, face=16, mips=5)
    assert sc.globals.reflection_capture_count == 70
    zero = np.zeros((sc.height, sc.width, 4), np.float32)
    got = run(ctx, sc, refl, zero, flags=abi.HL_REFLECTION_STORE)
    ref, _ = O.reflection(sc, refl, flags=abi.HL_REFLECTION_STORE)
    check_radiance(got, ref, sc.depth < 1.0)


def test_debug_modes_and_disabled(ctx):
    print("[reflection] debug modes 30, 31, 1 and everything disabled", flush=True)
    sc = synth.make_scene("c1_720p_64", width=640, height=360)
    refl = synth.make_reflections(sc, face=16, mips=5, rc=True)
    base = np.random.default_rng(4).uniform(0.0, 1.0, size=(sc.height, sc.width, 4)).astype(np.float32)
    g = abi.DeferredGlobals()
    C.memmove(C.byref(g), C.byref(sc.globals), C.sizeof(g))
    for mode in (30, 31, 1):
        g.debug_mode = mode
        got = run(ctx, sc, refl, base, globals_=g)
        ref, _ = O.reflection(sc, refl, out=base.copy(), globals_=g)
        assert np.array_equal(got, ref), mode
    g.debug_mode = 0
    g.enable_reflections = g.enable_env_reflections = g.has_rc_gi = 0
    got = run(ctx, sc, refl, base, globals_=g)
    assert np.array_equal(got, base)


def test_rejects_bad_inputs(ctx):
    sc = synth.make_scene("c1_720p_64", width=64, height=64)
    refl = synth.make_reflections(sc, face=8, mips=4)
    ds = upload_scene(ctx, sc)
    out = ctx.alloc(64 * 64 * 16)
    dr = DeviceReflections(ctx, refl)
    for field, value in (("mip_levels", 5), ("mip_levels", 0), ("face_size", 0), ("num_cubes", 0), ("format", 9)):
        saved = getattr(dr.inputs.env, field)
        setattr(dr.inputs.env, field, value)
        with pytest.raises(abi.HelioError):
            ReflectionPass(ctx).execute(ds, dr, out)
        setattr(dr.inputs.env, field, saved)
    g = abi.DeferredGlobals()
    C.memmove(C.byref(g), C.byref(sc.globals), C.sizeof(g))
    g.debug_mode = 3
    with pytest.raises(abi.HelioError):
        ReflectionPass(ctx).execute(ds, dr, out, globals_=g)
    ReflectionPass(ctx).execute(ds, dr, out)  # and the restored inputs run


@pytest.mark.parametrize("case", range(8))
def test_random_configurations(ctx, case):
    """Ragged frame sizes, 0-19 boxes + 1-19 spheres, cube faces of 8-64
    texels with any mip count, the three cube formats, RC on and off, added
    over a random frame; the scene (and so the G-buffer set: reference, fp16,
    fp32) and the cube format are drawn independently: within 1e-4 of the
    restatement (in practice a few 1e-7, from exp2 in env_brdf_approx)."""
    rng = np.random.default_rng(1000 + case)
    name = str(rng.choice(["c1_720p_64", "c2_1080p_1k", "c3_4k_16k", "c4_512_adj"]))
    over = dict(num_lights=1024) if name == "c3_4k_16k" else {}
    W, H = int(rng.integers(97, 700)), int(rng.integers(61, 400))
    sc = synth.make_scene(name, width=W, height=H, **over)
    face = int(rng.choice([8, 16, 32, 64]))
    mips = int(rng.integers(1, int(np.log2(face)) + 2))
    fmt = int(rng.choice([abi.HL_FMT_RGBA16_FLOAT, abi.HL_FMT_RGBA32_FLOAT, abi.HL_FMT_RGBA8_UNORM]))
    nb, ns = int(rng.integers(0, 20)), int(rng.integers(1, 20))
    print(f"[reflection] random case {case}: {name} {W}x{H}, {nb}+{ns} captures, face {face} x {mips} mips, "
          f"format {fmt}", flush=True)
    refl = synth.make_reflections(sc, n_box=nb, n_sphere=ns, face=face, mips=mips, fmt=fmt, rc=bool(case % 2),
                                  seed=int(rng.integers(1, 1000)))
    base = rng.uniform(0.0, 1.0, size=(H, W, 4)).astype(np.float32)
    got = run(ctx, sc, refl, base)
    ref, _ = O.reflection(sc, refl, out=base.copy())
    check_radiance(got, ref)
