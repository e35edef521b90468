"""Pins the oracle (CPU checker) without the reference, which cannot run here
(Rust + WGSL on wgpu; no cargo, no wgpu in this image).

* known answers derived from the formulas of pbr_direct_light / pbr_eval.wgsl
  (range cut-off, rough head-on closed form, metallic => no diffuse, spot cone
  limits, the 0.002 cut-off, directional lights, sky, ambient only);
* invariants (linearity in intensity and colour, list order, permutation);
* an independent FP64 geometric brute force of the tile cull (light_cull.wgsl)
  that may only disagree on pairs within 1e-5 of a plane;
* the FP64 implementation (separate code) against the FP32 restatement;
* conservativeness of the clustered lists by brute force over all lights.
"""
import ctypes as C
import math

import numpy as np
import pytest

from helio_amd import abi, synth
import oracle_lib as O

F32 = np.float32
PI32 = F32(3.14159265359)


def light(pos=(0, 0, 0), rng=10.0, color=(1, 1, 1), intensity=1.0, ltype=abi.HL_LIGHT_POINT,
          direction=(0, -1, 0), outer=0.0, inner=0.0):
    l = np.zeros(1, dtype=abi.GPU_LIGHT_DTYPE)
    l["position_range"] = (*pos, rng)
    l["direction_outer"] = (*direction, outer)
    l["color_intensity"] = (*color, intensity)
    l["shadow_index"] = abi.HL_NO_SHADOW
    l["light_type"] = ltype
    l["inner_angle"] = inner
    l["ies_profile_index"] = -1
    l["light_function_index"] = -1
    return l


def term(l, wp, N, V, F0, albedo, rough, metal):
    out = (C.c_float * 3)()
    O.oracle().or_pbr_direct_light(l.ctypes.data, O.f3(wp), O.f3(N), O.f3(V), O.f3(F0), O.f3(albedo),
                                   C.c_float(rough), C.c_float(metal), out)
    return np.array(out[:], dtype=np.float64)


def closed_form(rad, NdL, NdV, NdH, HdV, F0, albedo, rough, metal):
    """pbr_direct_light's value in float64 (deferred_lighting.wgsl:778-820 with
    pbr_eval.wgsl:12-56): (kD a/pi + D G F / (4 NdV NdL + 1e-4)) rad NdL."""
    pi = 3.14159265359
    a2 = rough ** 4
    D = a2 / (pi * (NdH * NdH * (a2 - 1) + 1) ** 2 + 1e-4)
    k = (rough + 1) ** 2 / 8
    g = lambda x: x / (x * (1 - k) + k + 1e-4)
    G = g(NdV) * g(NdL)
    P = (1 - HdV) ** 5
    F = np.asarray(F0, float) + (1 - np.asarray(F0, float)) * P
    kD = (1 - F) * (1 - metal)
    return (kD * np.asarray(albedo, float) / pi + D * G * F / (4 * NdV * NdL + 1e-4)) * np.asarray(rad, float) * NdL


def test_range_cutoff_is_zero():
    wp, N, V = (0, 0, 0), (0, 1, 0), (0, 1, 0)
    for d, r in ((5.0, 4.9), (4.0, 4.0)):  # beyond range; at range 1-(d/r)^4 = 0
        l = light(pos=(0, d, 0), rng=r, intensity=5.0)
        assert np.all(term(l, wp, N, V, (0.04,) * 3, (0.5,) * 3, 0.5, 0.0) == 0.0)


def test_rough_head_on_closed_form():
    # N = L = V, roughness 1: a2 = 1, D = 1/(pi + 1e-4), k = 0.5, F = F0
    d, I = 2.0, 3.0
    l = light(pos=(0, d, 0), rng=100.0, intensity=I)
    F0, alb = (0.04, 0.05, 0.06), (0.2, 0.4, 0.6)
    got = term(l, (0, 0, 0), (0, 1, 0), (0, 1, 0), F0, alb, 1.0, 0.0)
    atten = 1.0 / (d * d + 1e-4) * (1 - (d / 100.0) ** 4)
    rad = np.array([1.0, 1.0, 1.0]) * I * atten
    D = 1 / (3.14159265359 + 1e-4)
    G1 = 1 / (1 * 0.5 + 0.5 + 1e-4)
    spec = D * G1 * G1 * np.array(F0) / (4 + 1e-4)
    diff = (1 - np.array(F0)) * np.array(alb) / 3.14159265359
    want = (diff + spec) * rad
    np.testing.assert_allclose(got, want, rtol=2e-6)
    np.testing.assert_allclose(got, closed_form(rad, 1, 1, 1, 1, F0, alb, 1.0, 0.0), rtol=2e-6)


def test_metallic_removes_diffuse():
    l = light(pos=(1.0, 3.0, 0.5), rng=50.0, intensity=2.0)
    wp = (0.0, 0.0, 0.0)
    N = np.array([0.0, 1.0, 0.0])
    V = np.array([0.3, 0.9, 0.1]); V /= np.linalg.norm(V)
    L = np.array([1.0, 3.0, 0.5]); d = np.linalg.norm(L); L /= d
    H = (V + L) / np.linalg.norm(V + L)
    F0, alb, rough = (0.9, 0.6, 0.3), (0.9, 0.6, 0.3), 0.4
    got = term(l, wp, N, V, F0, alb, rough, 1.0)
    rad = np.array([2.0] * 3) / (d * d + 1e-4) * (1 - (d / 50.0) ** 4)
    want = closed_form(rad, N @ L, N @ V, N @ H, H @ V, F0, alb, rough, 1.0)
    np.testing.assert_allclose(got, want, rtol=2e-5)
    # with metal = 1 the albedo enters only through F0: changing it changes nothing
    assert np.array_equal(got, term(l, wp, N, V, F0, (0.0, 0.0, 0.0), rough, 1.0))


def test_spot_cone_limits():
    outer, inner = math.cos(math.radians(20)), math.cos(math.radians(16))
    base = dict(pos=(0, 4, 0), rng=20.0, intensity=3.0)
    wp, N, V = (0, 0, 0), (0, 1, 0), (0, 1, 0)
    args = ((0.04,) * 3, (0.5,) * 3, 0.5, 0.0)
    point = term(light(**base), wp, N, V, *args)
    # axis straight at the pixel: cos_a = 1 >= cos_inner -> factor 1, bit-identical to a point light
    on = term(light(**base, ltype=abi.HL_LIGHT_SPOT, direction=(0, -1, 0), outer=outer, inner=inner), wp, N, V, *args)
    assert np.array_equal(on, point) and np.all(point > 0)
    # axis 30 degrees away: cos_a <= cos_outer -> 0
    a = math.radians(30)
    off = term(light(**base, ltype=abi.HL_LIGHT_SPOT, direction=(math.sin(a), -math.cos(a), 0), outer=outer,
                     inner=inner), wp, N, V, *args)
    assert np.all(off == 0.0)
    # 18 degrees: between -> smoothstep in (0, 1)
    a = math.radians(18)
    mid = term(light(**base, ltype=abi.HL_LIGHT_SPOT, direction=(math.sin(a), -math.cos(a), 0), outer=outer,
                     inner=inner), wp, N, V, *args)
    t = (math.cos(a) - outer) / (inner - outer)
    np.testing.assert_allclose(mid, point * t * t * (3 - 2 * t), rtol=1e-4)


def test_radiance_cutoff_0002():
    # all(radiance < 0.002) -> exactly 0 (deferred_lighting.wgsl:781)
    wp, N, V = (0, 0, 0), (0, 1, 0), (0, 1, 0)
    args = ((0.04,) * 3, (0.5,) * 3, 0.5, 0.0)
    d = 3.0
    atten = 1.0 / (d * d + 1e-4) * (1 - (d / 1000.0) ** 4)
    I_lo = 0.00199 / atten
    I_hi = 0.00201 / atten
    assert np.all(term(light(pos=(0, d, 0), rng=1000.0, intensity=I_lo), wp, N, V, *args) == 0)
    assert np.all(term(light(pos=(0, d, 0), rng=1000.0, intensity=I_hi), wp, N, V, *args) > 0)
    # one channel above the cut-off keeps all three
    hi1 = term(light(pos=(0, d, 0), rng=1000.0, intensity=I_hi, color=(1, 0.5, 0.5)), wp, N, V, *args)
    assert np.all(hi1 > 0)


def test_directional_light():
    dirv = np.array([0.3, -1.0, 0.2]); dirv /= np.linalg.norm(dirv)
    l = light(ltype=abi.HL_LIGHT_DIRECTIONAL, direction=tuple(dirv), color=(1.0, 0.8, 0.6), intensity=2.0,
              pos=(1e6, 1e6, 1e6), rng=0.001)
    N = np.array([0.0, 1.0, 0.0]); V = np.array([0.0, 0.6, 0.8])
    L = -dirv
    H = (V + L) / np.linalg.norm(V + L)
    got = term(l, (5, 5, 5), N, V, (0.04,) * 3, (0.5, 0.5, 0.5), 0.6, 0.0)
    want = closed_form(np.array([1.0, 0.8, 0.6]) * 2.0, N @ L, N @ V, N @ H, H @ V, (0.04,) * 3, (0.5,) * 3, 0.6, 0.0)
    np.testing.assert_allclose(got, want, rtol=2e-5)


def test_facing_away_is_zero():
    l = light(pos=(0, -3, 0), rng=50.0, intensity=5.0)
    assert np.all(term(l, (0, 0, 0), (0, 1, 0), (0, 1, 0), (0.04,) * 3, (0.5,) * 3, 0.5, 0.0) == 0.0)


def test_linear_in_intensity_and_colour():
    wp, N = (0, 0, 0), (0.0, 1.0, 0.0)
    V = (0.0, 0.8, 0.6)
    args = ((0.04,) * 3, (0.3, 0.5, 0.7), 0.35, 0.0)
    a = term(light(pos=(1, 3, 2), rng=30, intensity=1.5, color=(0.7, 0.8, 0.9)), wp, N, V, *args)
    b = term(light(pos=(1, 3, 2), rng=30, intensity=3.0, color=(0.7, 0.8, 0.9)), wp, N, V, *args)
    c = term(light(pos=(1, 3, 2), rng=30, intensity=1.5, color=(1.4, 1.6, 1.8)), wp, N, V, *args)
    np.testing.assert_allclose(b, 2 * a, rtol=1e-6)
    np.testing.assert_allclose(c, 2 * a, rtol=1e-6)


def small_scene(**over):
    base = dict(width=160, height=96)
    base.update(over)
    return synth.make_scene("c2_1080p_1k", **base)


def test_sky_untouched_and_ambient_only():
    sc = synth.make_scene("c1_720p_64", width=96, height=64, num_lights=0)
    tiles = ((sc.width + 15) // 16) * ((sc.height + 15) // 16)
    lists = np.zeros(tiles * 64, np.uint32)
    counts = np.zeros(tiles, np.uint32)
    grid = O.tile_grid_host(sc, lists, counts)
    out = np.full((sc.height, sc.width, 4), 7.0, np.float32)
    O.shade(sc, grid, out=out)
    sky = sc.depth >= 1.0
    assert sky.any() and np.all(out[sky] == 7.0)
    O.shade(sc, grid, out=out, flags=abi.HL_SHADE_CLEAR_SKY)
    assert np.all(out[sky] == 0.0)
    # ambient only: mix(0.15 s, s, 0.5 N.y + 0.5) * albedo * ao + emissive
    alb = sc.albedo[..., :3].astype(np.float64) / 255.0
    ao = sc.orm[..., 0].astype(np.float64) / 255.0
    n = sc.normal.view(np.float16).reshape(sc.height, sc.width, 4)[..., :3].astype(np.float64)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    s = np.array([0.5, 0.5, 0.6])
    t = n[..., 1:2] * 0.5 + 0.5
    want = (0.15 * s * (1 - t) + s * t) * alb * ao[..., None]
    np.testing.assert_allclose(out[~sky][:, :3], want[~sky], rtol=1e-6, atol=1e-7)
    assert np.all(out[~sky][:, 3] == 1.0)


def fp64_tile_hits(sc):
    """Independent FP64 tile/sphere test (light_cull.wgsl:85-117 in exact-ish arithmetic)."""
    cam = sc.camera
    p00, p11 = float(cam.proj[0]), float(cam.proj[5])
    view = np.array(cam.view[:], dtype=np.float64).reshape(4, 4).T  # row-major
    W, H = sc.width, sc.height
    ntx, nty = (W + 15) // 16, (H + 15) // 16
    pos = np.c_[sc.lights["position_range"][:, :3].astype(np.float64), np.ones(len(sc.lights))]
    c = (view @ pos.T).T[:, :3]
    r = sc.lights["position_range"][:, 3].astype(np.float64)
    hits, margin = [], []
    for ty in range(nty):
        for tx in range(ntx):
            l = (tx * 16 / W) * 2 - 1
            rr = min(((tx + 1) * 16 / W) * 2 - 1, 1.0)
            t = 1 - (ty * 16 / H) * 2
            b = max(1 - ((ty + 1) * 16 / H) * 2, -1.0)

            def plane(p0, p1):
                r0 = np.array([p0[0] / p00, p0[1] / p11, -1.0]); r0 /= np.linalg.norm(r0)
                r1 = np.array([p1[0] / p00, p1[1] / p11, -1.0]); r1 /= np.linalg.norm(r1)
                n = np.cross(r0, r1)
                return n / np.linalg.norm(n)
            P = np.stack([plane((l, b), (l, t)), plane((rr, t), (rr, b)), plane((l, t), (rr, t)),
                          plane((rr, b), (l, b))])
            dist = c @ P.T  # [L, 4]
            inside = np.all(dist >= -r[:, None], axis=1)
            near = np.any(np.abs(dist + r[:, None]) < 1e-5 * (np.linalg.norm(c, axis=1) + r)[:, None], axis=1)
            hits.append(inside)
            margin.append(near)
    return np.array(hits), np.array(margin)


@pytest.mark.parametrize("seed_off", [0, 1, 2])
def test_tile_cull_against_fp64_geometry(seed_off):
    sc = small_scene(seed=synth.SEED_BASE + 977 * (seed_off + 1), num_lights=200)
    lists, counts = O.cull_tiles(sc)
    hits, margin = fp64_tile_hits(sc)
    for t in range(len(counts)):
        got = lists[t, :counts[t]]
        assert np.all(np.diff(got.astype(np.int64)) > 0), "ascending, unique"
        want = np.nonzero(hits[t])[0]
        assert counts[t] == min(len(want), 64)
        # identical up to lights within 1e-5 of a plane
        diff = set(got.tolist()) ^ set(want[:64].tolist())
        if counts[t] < 64:
            assert all(margin[t, i] for i in diff), (t, diff)


def test_tile_cull_cap_keeps_lowest_and_directional_everywhere():
    sc = small_scene(num_lights=300, seed=synth.SEED_BASE + 5)
    lights = sc.lights.copy()
    lights[7]["light_type"] = abi.HL_LIGHT_DIRECTIONAL
    lights[7]["position_range"] = (1e9, 1e9, 1e9, 0.0)
    sc.lights = lights
    lists, counts = O.cull_tiles(sc)
    hits, _ = fp64_tile_hits(sc)
    for t in range(len(counts)):
        assert 7 in lists[t, :counts[t]] or counts[t] == 64 and lists[t, 63] < 7
        if counts[t] == 64:
            assert np.all(np.diff(lists[t].astype(np.int64)) > 0)


def test_tile_cull_permutation_invariance():
    sc = small_scene(num_lights=40, seed=synth.SEED_BASE + 11)
    l1, c1 = O.cull_tiles(sc)
    perm = np.random.default_rng(3).permutation(len(sc.lights))
    sc2 = small_scene(num_lights=40, seed=synth.SEED_BASE + 11)
    sc2.lights = sc.lights[perm].copy()
    l2, c2 = O.cull_tiles(sc2)
    assert np.array_equal(c1, c2)
    for t in range(len(c1)):
        assert set(l1[t, :c1[t]].tolist()) == set(perm[l2[t, :c2[t]]].tolist())


@pytest.mark.parametrize("name,over", [("c1_720p_64", dict(width=320, height=180)),
                                       ("c2_1080p_1k", dict(width=480, height=270)),
                                       ("c3_4k_16k", dict(width=640, height=360, num_lights=4096)),
                                       ("c4_512_adj", dict(width=256, height=256))])
def test_cluster_lists_are_conservative(hlib, name, over):
    """Every (pixel, light) pair with a nonzero term is in the pixel's cluster."""
    sc = synth.make_scene(name, **over)
    p = synth.cluster_params(hlib, sc, cap=0)
    off, cnt, pool, flags = O.cull_clusters(sc, p, 1 << 22)
    assert flags == 0
    grid = O.cluster_grid_host(p, off, cnt, pool)
    seen = C.c_uint64()
    bad = O.oracle().or_check_conservative(C.byref(sc.camera), C.byref(sc.globals), C.byref(sc.gbuffer_host()),
                                           sc.lights.ctypes.data, len(sc.lights), C.byref(grid), 0, 2,
                                           C.byref(seen))
    assert seen.value > 1000
    assert bad == 0
    # and with clusters, shading equals shading over all lights (no cap)
    out_c, q_c = O.shade(sc, grid)
    n = len(sc.lights)
    allc = np.full(p.grid_x * p.grid_y * p.grid_z, n, np.uint32)
    alloff = np.zeros_like(allc)
    allpool = np.arange(n, dtype=np.uint32)
    out_a, q_a = O.shade(sc, O.cluster_grid_host(p, alloff, allc, allpool))
    assert np.array_equal(out_c, out_a)


def test_cluster_cap_is_prefix_of_uncapped(hlib):
    sc = synth.make_scene("c2_1080p_1k", width=480, height=270)
    p0 = synth.cluster_params(hlib, sc, cap=0)
    p8 = synth.cluster_params(hlib, sc, cap=8)
    o0, c0, pool0, f0 = O.cull_clusters(sc, p0, 1 << 22)
    o8, c8, pool8, f8 = O.cull_clusters(sc, p8, 1 << 22)
    assert f0 == 0 and (f8 & abi.HL_OVERFLOW_CAP)
    assert np.array_equal(c8, np.minimum(c0, 8))
    for c in range(len(c0)):
        a = pool0[o0[c]:o0[c] + c0[c]]
        b = pool8[o8[c]:o8[c] + c8[c]]
        assert np.array_equal(b, a[:8])
        assert np.all(np.diff(a.astype(np.int64)) > 0)
    assert np.array_equal(o8[1:], np.cumsum(c8)[:-1])
    # pool truncation: nothing past pool_cap, counts shrink, flag set
    cap = int(c8.sum()) // 2
    ot, ct, pt, ft = O.cull_clusters(sc, p8, cap)
    assert ft & abi.HL_OVERFLOW_POOL and int(ct.sum()) == cap
    assert np.all(ot + ct <= cap)


def test_slice_rule_matches_thresholds(hlib):
    sc = synth.make_scene("c3_4k_16k", width=320, height=180, num_lights=16)
    p = synth.cluster_params(hlib, sc)
    orc = O.oracle()
    for k in range(1, p.grid_z):
        d = p.depth_split[k]
        assert orc.or_slice_of_depth(C.byref(p), C.c_float(d)) == k
        assert orc.or_slice_of_depth(C.byref(p), C.c_float(np.nextafter(F32(d), F32(0)))) == k - 1
    assert orc.or_slice_of_depth(C.byref(p), C.c_float(0.0)) == 0
    assert orc.or_slice_of_depth(C.byref(p), C.c_float(0.99999994)) == p.grid_z - 1


@pytest.mark.parametrize("name,over", [("c4_512_adj", dict(width=128, height=128)),
                                       ("c2_1080p_1k", dict(width=192, height=108, gbuf_format=synth.GBUF_F32))])
def test_fp32_restatement_against_fp64(name, over):
    """The FP32 restatement and the separately written FP64 forward agree to
    FP32 rounding.  Both start from the same FP32 world position; an fp32
    G-buffer stores F0 = mix(0.04, albedo, metallic) from the same albedo and
    metallic the FP64 path resolves it from.  A value may differ only where a
    term sits within rounding of a cut-off (range, 0.002) -- at most 0.1 %."""
    sc = synth.make_scene(name, **over)
    n = len(sc.lights)
    W, H = sc.width, sc.height
    allc = np.full(1, n, np.uint32)
    p = abi.hl_cluster_params()
    p.grid_x = p.grid_y = p.grid_z = 1
    p.cell_w, p.cell_h, p.width, p.height = W, H, W, H
    off0, pool = np.zeros(1, np.uint32), np.arange(n, dtype=np.uint32)
    grid = O.cluster_grid_host(p, off0, allc, pool)  # grid holds raw pointers: keep the arrays alive
    out32, _ = O.shade(sc, grid)
    px = (O.or64_pixel * (W * H))()
    O.oracle().or64_decode(C.byref(sc.camera), C.byref(sc.gbuffer_host()), px)
    out64 = np.zeros((H, W, 3))
    O.oracle().or64_forward(px, W, H, C.byref(sc.globals), sc.lights.ctypes.data, n,
                            out64.ctypes.data, None, 0)
    m = sc.depth < 1.0
    err = np.abs(out32[..., :3] - out64) / (np.abs(out64) + 1e-6)
    frac_bad = float((err[m] > 1e-4).mean())
    assert frac_bad < 1e-3, frac_bad
    assert float(np.median(err[m])) < 1e-6
