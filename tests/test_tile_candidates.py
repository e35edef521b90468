"""The binned tile cull's candidate tiles contain every member.

hl_light_cull_tiles bins each light into the super-tiles (4 x 4 tiles) that
its candidate tiles touch (light_tiles in hl_cull.hip), and each tile tests
only its super-tile's candidates, in ascending order, with the exact plane
test of light_cull.wgsl.  The candidates come from necessary conditions of
passing a tile's four planes: in front of the camera the projected bound of
the cluster cull; behind it (a tile's planes through the origin meet only
near the mirrored projection) l >= p00 (-cx - r fx) / |zc|, r <= p00 (r fx -
cx) / |zc| and alike for rows; every tile within 1e-4 of the camera plane.
Restated here in float32; every entry of the oracle's tile lists (scenes of
at most 64 lights, so that no list is truncated) must lie in its light's
rectangle, for the 64-light scene and for lights around, behind and beside
the camera with radii up to 30.  Planted mistakes (the front bound used
behind the camera, untilted planes) miss members.
"""
import numpy as np
import pytest

import oracle_lib as O
from helio_amd import abi, synth

f32 = np.float32


def light_tiles(sc, mistake=None):
    """[n] ok, tx0, tx1, ty0, ty1 of every light (float32, as the kernel)."""
    from helio_amd.device import light_cull_params
    p = light_cull_params(sc)
    ntx, nty = p.num_tiles_x, p.num_tiles_y
    cam = sc.camera
    view = np.array(cam.view, f32).reshape(4, 4)  # view[col][row]
    L = sc.lights
    pos = L["position_range"][:, :3].astype(f32)
    vp = np.stack([((view[0, r] * pos[:, 0] + view[1, r] * pos[:, 1]) + view[2, r] * pos[:, 2]) + view[3, r]
                   for r in range(3)], 1).astype(f32)
    rng = np.where(L["light_type"] == abi.HL_LIGHT_DIRECTIONAL, f32(np.inf), L["position_range"][:, 3]).astype(f32)
    n = len(L)
    zc = (-vp[:, 2]).astype(f32)
    p00, p11 = f32(cam.proj[0]), f32(cam.proj[5])
    full = ~np.isfinite(rng) | ~np.isfinite(vp).all(1) | (np.abs(zc) <= f32(1e-4)) | ~((p00 > 0) & (p11 > 0))
    with np.errstate(all="ignore"):
        slack = (f32(1e-5) * ((np.abs(vp[:, 0]) + np.abs(vp[:, 1])) + np.abs(zc))).astype(f32)
        rr = (np.abs(rng) * f32(1 + 1e-5) + slack).astype(f32)
        fx = np.sqrt(f32(1) + f32(1) / (p00 * p00))
        fy = np.sqrt(f32(1) + f32(1) / (p11 * p11))
        if mistake == "no_tilt":
            fx = fy = f32(1)
        iz = f32(1) / np.abs(zc)
        cx, cy = vp[:, 0], vp[:, 1]
        front = (zc > 0) | (mistake == "front_only")
        nx0 = np.where(front, p00 * (cx - rr * fx) * iz, p00 * (-cx - rr * fx) * iz)
        nx1 = np.where(front, p00 * (cx + rr * fx) * iz, p00 * (rr * fx - cx) * iz)
        nb = np.where(front, p11 * (cy - rr * fy) * iz, p11 * (-rr * fy - cy) * iz)
        nt = np.where(front, p11 * (cy + rr * fy) * iz, p11 * (rr * fy - cy) * iz)
        fw, fh = f32(p.screen_width), f32(p.screen_height)
        px0, px1 = (nx0 + 1) * f32(0.5) * fw, (nx1 + 1) * f32(0.5) * fw
        py0, py1 = (1 - nt) * f32(0.5) * fh, (1 - nb) * f32(0.5) * fh
        px0 = px0 - (f32(1e-5) * np.abs(px0) + f32(0.25))
        px1 = px1 + (f32(1e-5) * np.abs(px1) + f32(0.25))
        py0 = py0 - (f32(1e-5) * np.abs(py0) + f32(0.25))
        py1 = py1 + (f32(1e-5) * np.abs(py1) + f32(0.25))
        its = f32(1 / 16)
        x0 = np.fmin(np.fmax(np.floor(px0 * its), -1), ntx)
        x1 = np.fmax(np.fmin(np.floor(px1 * its), ntx), -1)
        y0 = np.fmin(np.fmax(np.floor(py0 * its), -1), nty)
        y1 = np.fmax(np.fmin(np.floor(py1 * its), nty), -1)
    x0 = np.where(full, 0, np.clip(x0, 0, ntx - 1)).astype(np.int64)
    x1 = np.where(full, ntx - 1, np.clip(x1, -1, ntx - 1)).astype(np.int64)
    y0 = np.where(full, 0, np.clip(y0, 0, nty - 1)).astype(np.int64)
    y1 = np.where(full, nty - 1, np.clip(y1, -1, nty - 1)).astype(np.int64)
    return x0, x1, y0, y1, ntx


def misses(sc, mistake=None):
    lists, counts = O.cull_tiles(sc)
    assert counts.max() < 64 or len(sc.lights) <= 64
    x0, x1, y0, y1, ntx = light_tiles(sc, mistake)
    tiles = np.repeat(np.arange(len(counts)), counts.astype(np.int64))
    li = np.concatenate([lists[t, :c] for t, c in enumerate(counts)]).astype(np.int64)
    tx, ty = tiles % ntx, tiles // ntx
    inside = (tx >= x0[li]) & (tx <= x1[li]) & (ty >= y0[li]) & (ty <= y1[li])
    return int((~inside).sum()), len(li)


def around_camera(seed, n=64, width=640, height=360):
    """Point lights around the camera: in front, beside, behind (to 20 m)
    and within 1e-4 of its plane, radii 0.5 to 30."""
    sc = synth.make_scene("c1_720p_64", width=width, height=height, num_lights=n, seed=synth.SEED_BASE + seed)
    rng = np.random.default_rng(seed)
    cam = sc.camera
    inv = np.linalg.inv(np.array(cam.view, np.float64).reshape(4, 4).T)
    zc = np.concatenate([rng.uniform(-20.0, -0.01, n // 2), rng.uniform(0.01, 40.0, n // 4),
                         rng.uniform(-1e-4, 1e-4, n - n // 2 - n // 4)])
    v = np.c_[rng.uniform(-25, 25, n), rng.uniform(-15, 15, n), -zc, np.ones(n)]
    w = (inv @ v.T).T[:, :3]
    L = sc.lights
    L["light_type"] = abi.HL_LIGHT_POINT
    L["position_range"][:, :3] = w
    L["position_range"][:, 3] = rng.uniform(0.5, 30.0, n)
    return sc


@pytest.mark.parametrize("seed", range(6))
def test_candidates_hold_every_member(seed):
    sc = synth.make_scene("c1_720p_64", width=640, height=360, seed=synth.SEED_BASE + 100 + seed)
    miss, members = misses(sc)
    assert members > 200 and miss == 0, (miss, members)
    sc = around_camera(seed)
    miss, members = misses(sc)
    assert members > 200 and miss == 0, (miss, members)


@pytest.mark.parametrize("mistake", ["front_only", "no_tilt"])
def test_planted_mistakes_miss_members(mistake):
    missed = sum(misses(around_camera(seed), mistake)[0] for seed in range(6))
    assert missed > 0, mistake
