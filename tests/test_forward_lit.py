"""Forward-lit fragment shading (forward_lit.wgsl fs_main), CPU side.

The forward oracle (oracle/forward_lit.c) is pinned by
* the light direction of a Transform against rotation matrices in double
  (an independent construction: R = Ry(yaw) Rx(pitch) Rz(roll) applied to -Z)
  and exact axis cases;
* the per-light term against the deferred oracle's pbr_direct_light, which
  the known answers and the FP64 forward pin: bit-equal wherever the deferred
  shader's 0.002 cut-off does not apply, and nonzero only below it otherwise;
* closed forms: the rough head-on light, ambient + emissive with no light,
  the discard threshold, a light that only the forward shader keeps;
* the light indirection through entity indices and transforms, equal bit for
  bit to lights that carry the transform's position and direction;
* brute force: shading with the tile or cluster lists equals shading with
  every light listed, bit for bit (the lists are conservative for fragments).
The C ABI's argument checks run without a GPU.
"""
import ctypes as C
import math

import numpy as np
import pytest

from helio_amd import abi, synth
import oracle_lib as O
from test_oracle import closed_form, light, term

F32 = np.float32


def direction(rot):
    out = (C.c_float * 3)()
    O.oracle().or_light_direction_from_rotation((C.c_float * 3)(*[float(v) for v in rot]), out)
    return np.array(out[:], dtype=np.float32)


def fwd_term(l, pos, d, wp, N, V, F0, albedo, rough, metal):
    out = (C.c_float * 3)()
    O.oracle().or_forward_pbr_direct_light(l.ctypes.data, O.f3(pos), O.f3(d), O.f3(wp), O.f3(N), O.f3(V),
                                           O.f3(F0), O.f3(albedo), C.c_float(rough), C.c_float(metal), out)
    return np.array(out[:], dtype=np.float32)


# ------------------------------------------------------------- direction
def test_direction_axis_cases():
    assert np.array_equal(direction((0, 0, 0)), F32([0, 0, -1]))
    for roll in (17.0, -90.0, 180.0):  # a roll turns -Z about Z: unchanged
        assert np.array_equal(direction((0, 0, roll)), F32([0, 0, -1]))
    np.testing.assert_allclose(direction((0, 90, 0)), [-1, 0, 0], atol=1e-7)   # yaw: -Z -> -X
    np.testing.assert_allclose(direction((90, 0, 0)), [0, 1, 0], atol=1e-7)    # pitch: -Z -> +Y
    np.testing.assert_allclose(direction((0, 180, 0)), [0, 0, 1], atol=1e-7)


def test_direction_matches_rotation_matrices():
    rng = np.random.default_rng(3)
    rot = rng.uniform(-360, 360, size=(2000, 3)).astype(np.float32)
    got = np.array([direction(r) for r in rot])
    want = synth.direction_from_rotation_f64(rot)
    np.testing.assert_allclose(got, want, atol=3e-6)
    np.testing.assert_allclose(np.linalg.norm(got.astype(np.float64), axis=1), 1.0, atol=2e-7)


# ------------------------------------------------------- per-light term
def test_forward_term_equals_deferred_term_above_cutoff():
    rng = np.random.default_rng(5)
    same = cut = 0
    for i in range(3000):
        ltype = (abi.HL_LIGHT_DIRECTIONAL, abi.HL_LIGHT_POINT, abi.HL_LIGHT_SPOT)[i % 3]
        dvec = rng.normal(size=3)
        dvec /= np.linalg.norm(dvec)
        outer = math.cos(math.radians(rng.uniform(10, 60)))
        l = light(pos=tuple(rng.uniform(-4, 4, 3)), rng=float(rng.uniform(1, 8)), color=tuple(rng.uniform(0, 1, 3)),
                  intensity=float(10 ** rng.uniform(-3, 1)), ltype=ltype, direction=tuple(dvec), outer=outer,
                  inner=min(1.0, outer + 0.02))
        N = rng.normal(size=3); N /= np.linalg.norm(N)
        V = rng.normal(size=3); V /= np.linalg.norm(V)
        args = (tuple(rng.uniform(-1, 1, 3)), tuple(N), tuple(V), tuple(rng.uniform(0, 0.999, 3)),
                tuple(rng.uniform(0, 1, 3)), float(rng.uniform(0.045, 1)), float(rng.uniform(0, 1)))
        d = term(l, *args).astype(np.float32)
        f = fwd_term(l, l["position_range"][0, :3], l["direction_outer"][0, :3], *args)
        if np.any(d != 0):
            assert np.array_equal(d, f), i
            same += 1
        elif np.any(f != 0):
            # only the deferred shader's cut-off can have removed it: every
            # channel of the radiance below 0.002
            ci = l["color_intensity"][0]
            if ltype == abi.HL_LIGHT_DIRECTIONAL:
                rad = ci[:3] * ci[3]
            else:
                wp = np.asarray(args[0], np.float64)
                dist = np.linalg.norm(l["position_range"][0, :3] - wp)
                att = 1 / (dist * dist + 1e-4) * max(0.0, 1 - (dist / l["position_range"][0, 3]) ** 4)
                if ltype == abi.HL_LIGHT_SPOT:
                    Lv = (l["position_range"][0, :3] - wp) / dist
                    e0, e1 = l["direction_outer"][0, 3], l["inner_angle"][0]
                    tt = np.clip((np.dot(-Lv, l["direction_outer"][0, :3]) - e0) / (e1 - e0), 0, 1)
                    att *= tt * tt * (3 - 2 * tt)
                rad = ci[:3] * ci[3] * att
            assert np.all(rad < 0.002 * (1 + 1e-4)), (i, rad)
            cut += 1
    assert same > 300 and cut > 30, (same, cut)


def test_forward_keeps_a_light_below_the_deferred_cutoff():
    d = 3.0
    atten = 1.0 / (d * d + 1e-4) * (1 - (d / 1000.0) ** 4)
    l = light(pos=(0, d, 0), rng=1000.0, intensity=0.0015 / atten)
    args = ((0, 0, 0), (0, 1, 0), (0, 1, 0), (0.04,) * 3, (0.5,) * 3, 1.0, 0.0)
    assert np.all(term(l, *args) == 0)
    got = fwd_term(l, (0, d, 0), (0, -1, 0), *args)
    want = closed_form(np.full(3, 0.0015), 1, 1, 1, 1, (0.04,) * 3, (0.5,) * 3, 1.0, 0.0)
    np.testing.assert_allclose(got, want, rtol=5e-6)


# ------------------------------------------------------------ fragments
def frag(wp=(0, 0, 0), pixel=(0, 0), N=(0, 1, 0), alpha=1.0, albedo=(0.2, 0.4, 0.6), ao=1.0, f0=(0.04,) * 3,
         rough=1.0, emissive=(0, 0, 0), metal=0.0):
    f = np.zeros(1, dtype=abi.FRAGMENT_DTYPE)
    f["world_pos"], f["pixel"], f["normal"], f["alpha"] = wp, pixel[0] | (pixel[1] << 16), N, alpha
    f["albedo"], f["ao"], f["specular_f0"], f["roughness"] = albedo, ao, f0, rough
    f["emissive"], f["metallic"] = emissive, metal
    return f


def camera_at(eye):
    cam = abi.GpuCameraUniforms()
    cam.position_near[:] = (*eye, 0.1)
    # identity view_proj: depth = z (only cluster mode reads it)
    for k in range(4):
        cam.view_proj[k * 4 + k] = 1.0
    return cam


def fglobals(w=32, h=32, amb=(0.1, 0.2, 0.3), ai=0.5):
    g = abi.ForwardLitGlobals()
    g.ambient_intensity = ai
    g.ambient_color[:] = (*amb, 1.0)
    g.num_tiles_x, g.num_tiles_y = (w + 15) // 16, (h + 15) // 16
    g.screen_width, g.screen_height = float(w), float(h)
    return g


def one_tile_grid(indices, w=32, h=32):
    ntx, nty = (w + 15) // 16, (h + 15) // 16
    lists = np.zeros(ntx * nty * 64, np.uint32)
    counts = np.zeros(ntx * nty, np.uint32)
    for t in range(ntx * nty):
        lists[t * 64: t * 64 + len(indices)] = indices
        counts[t] = len(indices)
    g = O.tile_grid_host(type("S", (), {"width": w, "height": h})(), lists, counts)
    g._arrays = (lists, counts)
    return g


def test_head_on_fragment_closed_form():
    d, I = 2.0, 3.0
    lights = light(pos=(0, d, 0), rng=100.0, intensity=I)
    f = frag(wp=(0, 0, 0), N=(0, 1, 0), albedo=(0.2, 0.4, 0.6), f0=(0.04, 0.05, 0.06), rough=1.0)
    g = fglobals(amb=(0, 0, 0))
    out, pairs = O.forward_lit(camera_at((0, 5, 0)), g, f, lights, one_tile_grid([0]))
    atten = 1.0 / (d * d + 1e-4) * (1 - (d / 100.0) ** 4)
    want = closed_form(np.full(3, I * atten), 1, 1, 1, 1, (0.04, 0.05, 0.06), (0.2, 0.4, 0.6), 1.0, 0.0)
    np.testing.assert_allclose(out[0, :3], want, rtol=2e-6)
    assert out[0, 3] == 1.0 and pairs == 1


def test_no_light_is_ambient_plus_emissive():
    rng = np.random.default_rng(2)
    fr = np.concatenate([frag(albedo=tuple(rng.uniform(0, 1, 3)), ao=float(rng.uniform()),
                              emissive=tuple(rng.uniform(0, 2, 3)), alpha=float(rng.uniform(0.01, 1)),
                              pixel=(int(rng.integers(0, 32)), int(rng.integers(0, 32)))) for _ in range(50)])
    g = fglobals(amb=(0.3, 0.6, 0.9), ai=0.7)
    out, pairs = O.forward_lit(camera_at((0, 5, 0)), g, fr, np.zeros(0, abi.GPU_LIGHT_DTYPE), one_tile_grid([]))
    amb = np.asarray(g.ambient_color[:3], F32) * F32(g.ambient_intensity)
    want = (F32(0) + amb * fr["albedo"] * fr["ao"][:, None]) + fr["emissive"]
    assert np.array_equal(out[:, :3], want) and np.array_equal(out[:, 3], fr["alpha"]) and pairs == 0


def test_discard_threshold():
    lights = light(pos=(0, 2, 0), rng=100.0, intensity=3.0)
    a_keep = float(np.nextafter(F32(0.001), F32(1)))
    fr = np.concatenate([frag(alpha=0.001), frag(alpha=a_keep), frag(alpha=0.0), frag(alpha=-1.0),
                         frag(alpha=float("nan"))])
    out, pairs = O.forward_lit(camera_at((0, 5, 0)), fglobals(), fr, lights, one_tile_grid([0]))
    assert np.all(out[0] == 0) and np.all(out[2] == 0) and np.all(out[3] == 0)
    assert out[1, 3] == F32(a_keep) and np.all(out[1, :3] > 0)
    assert np.isnan(out[4, 3]) and np.all(out[4, :3] > 0)   # NaN <= 0.001 is false: kept
    assert pairs == 2


def test_pixel_outside_the_grid_gets_no_light():
    lights = light(pos=(0, 2, 0), rng=100.0, intensity=3.0)
    fr = np.concatenate([frag(pixel=(5, 5)), frag(pixel=(32, 5)), frag(pixel=(5, 40000))])
    out, pairs = O.forward_lit(camera_at((0, 5, 0)), fglobals(), fr, lights, one_tile_grid([0]))
    amb_only = O.forward_lit(camera_at((0, 5, 0)), fglobals(), fr, np.zeros(0, abi.GPU_LIGHT_DTYPE),
                             one_tile_grid([]))[0]
    assert np.all(out[0, :3] > amb_only[0, :3])
    assert np.array_equal(out[1:], amb_only[1:]) and pairs == 1


def scene_with_transforms(name="c2_1080p_1k", **over):
    sc = synth.make_scene(name, **over)
    lights, entity, tr = synth.make_light_transforms(sc.lights, seed=4)
    return sc, lights, entity, tr


def test_entity_indirection_equals_lights_carrying_the_transforms():
    sc, lights, entity, tr = scene_with_transforms(width=192, height=108)
    fr = synth.make_fragments(sc)
    g = synth.forward_globals(sc)
    lists, counts = O.cull_tiles(type("S", (), {"camera": sc.camera, "lights": lights, "width": sc.width,
                                                "height": sc.height, "globals": sc.globals})())
    grid = O.tile_grid_host(sc, lists.ravel(), counts)
    # the engine's placeholder: zeroed positions and a fixed direction in the light records
    placeholder = lights.copy()
    placeholder["position_range"][:, :3] = 0.0
    placeholder["direction_outer"][:, :3] = (0.0, -1.0, 0.0)
    via, pairs_via = O.forward_lit(sc.camera, g, fr, placeholder, grid, entity, tr)
    carried = lights.copy()
    carried["position_range"][:, :3] = tr["position"][entity]
    carried["direction_outer"][:, :3] = [direction(r) for r in tr["rotation"][entity]]
    own, pairs_own = O.forward_lit(sc.camera, g, fr, carried, grid)
    assert np.array_equal(via, own) and pairs_via == pairs_own > 0


def test_out_of_range_entity_contributes_nothing():
    sc, lights, entity, tr = scene_with_transforms(width=192, height=108)
    fr = synth.make_fragments(sc)
    g = synth.forward_globals(sc)
    grid = O.all_lights_grid(len(lights), sc.width, sc.height)
    bad = entity.copy()
    drop = np.arange(0, len(lights), 7)
    bad[drop] = len(tr) + drop
    got, _ = O.forward_lit(sc.camera, g, fr, lights, grid, bad, tr)
    keep = np.setdiff1d(np.arange(len(lights)), drop).astype(np.uint32)
    g2 = O.cluster_grid_host(grid._params, np.zeros(1, np.uint32), np.array([len(keep)], np.uint32), keep)
    want, _ = O.forward_lit(sc.camera, g, fr, lights, g2, entity, tr)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", ["tiles", "clusters"])
def test_lists_are_conservative_for_fragments(hlib, mode):
    """Shading with the cell lists equals shading with every light listed:
    a light missing from a fragment's list would change its sum."""
    sc, lights, entity, tr = scene_with_transforms("c1_720p_64", width=320, height=180)
    fr = synth.make_fragments(sc)
    g = synth.forward_globals(sc)
    if mode == "tiles":
        lists, counts = O.cull_tiles(type("S", (), {"camera": sc.camera, "lights": lights, "width": sc.width,
                                                    "height": sc.height, "globals": sc.globals})())
        assert counts.max() < 64  # uncapped, else the reference's own truncation applies
        grid = O.tile_grid_host(sc, lists.ravel(), counts)
    else:
        params = synth.cluster_params(hlib, sc)
        off, cnt, pool, flags = O.cull_clusters(type("S", (), {"camera": sc.camera, "lights": lights})(), params,
                                                1 << 20)
        assert flags == 0
        grid = O.cluster_grid_host(params, off, cnt, pool)
    got, _ = O.forward_lit(sc.camera, g, fr, lights, grid, entity, tr)
    want, _ = O.forward_lit(sc.camera, g, fr, lights, O.all_lights_grid(len(lights), sc.width, sc.height),
                            entity, tr)
    assert np.array_equal(got, want), np.count_nonzero(np.any(got != want, axis=1))
    dark, _ = O.forward_lit(sc.camera, g, fr, lights[:0], grid)
    assert np.mean(np.any(got != dark, axis=1)) > 0.3  # most fragments are lit


# ------------------------------------------------------------- C ABI
def test_forward_lit_rejects_bad_arguments_without_a_gpu(hlib):
    g = abi.ForwardLitGlobals()
    grid = abi.hl_light_grid()
    # a null context is refused before anything touches a device
    assert hlib.raw["hl_forward_lit"](None, None, C.byref(g), None, 0, None, 0, None, None, 0, C.byref(grid),
                                      None, 0) == abi.HL_EINVAL
    n = C.c_uint64()
    assert hlib.raw["hl_forward_lit_count_pairs"](None, None, C.byref(g), None, 0, None, 0, None, None, 0,
                                                  C.byref(grid), C.byref(n)) == abi.HL_EINVAL
