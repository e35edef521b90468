"""The clustered cull's candidate cells contain every member.

k_cluster_mark tests only the cells of each light's candidate box
(light_cells in hl_cull.hip): the columns and slices that necessary
conditions of the membership rule leave -- the box test's depth range
[slice_near - rc, slice_far + rc] around the light and, for each lateral
plane, l <= p00 (cx + rc sqrt(1 + l^2/p00^2)) / zc <= p00 (cx + rc fx) / zc --
widened by rounding margins.  batch_emu.light_boxes restates that computation
in float32.  Every (cell, light) pair of the oracle's lists (each light
tested against each cell, oracle.c or_light_cull_clusters) must lie in the
light's box: on the benchmark scenes at reduced size and on lights placed a
few ulp from slice limits and column planes, where the FP32 tests decide by
rounding.  Planted mistakes in the restatement (no slice padding, untilted
planes, the padding of the wrong slice) miss members there.
"""
import numpy as np
import pytest

import batch_emu as E
import oracle_lib as O
from helio_amd import abi, synth

f32 = np.float32


@pytest.fixture(scope="module")
def lib():
    return abi.Lib()


def boxes_of(sc, p, mistake=None):
    v = E.EmuView(sc.width, sc.height, sc.camera, sc.lights, np.arange(len(sc.lights)), False, p)
    return E.light_boxes(v, mistake)


def misses(sc, p, mistake=None):
    """Members of the oracle's lists outside their light's candidate box."""
    off, cnt, pool, flags = O.cull_clusters(sc, p, 1 << 24)
    assert flags == 0
    ok, x0, y0, s0, nx, ny, nz = boxes_of(sc, p, mistake)
    cells = np.repeat(np.arange(len(cnt)), cnt.astype(np.int64))
    li = np.concatenate([pool[o:o + n] for o, n in zip(off, cnt)]).astype(np.int64) if cnt.sum() else \
        np.zeros(0, np.int64)
    ncol = p.grid_x * p.grid_y
    k, col = cells // ncol, cells % ncol
    cx, cy = col % p.grid_x, col // p.grid_x
    inside = ok[li] & (cx >= x0[li]) & (cx < x0[li] + nx[li]) & (cy >= y0[li]) & (cy < y0[li] + ny[li]) & \
        (k >= s0[li]) & (k < s0[li] + nz[li])
    directional = sc.lights["light_type"][li] == abi.HL_LIGHT_DIRECTIONAL
    return int((~inside & ~directional).sum()), len(li)


@pytest.mark.parametrize("name,over,grid", [
    ("c1_720p_64", dict(width=320, height=180), None),
    ("c2_1080p_1k", dict(width=480, height=270), (16, 9, 24)),
    ("c3_4k_16k", dict(width=960, height=540, num_lights=4096), (32, 18, 48)),
    ("c3_4k_16k_far", dict(width=960, height=540, num_lights=4096), (32, 18, 48)),
])
def test_candidates_hold_every_member(lib, name, over, grid):
    sc = synth.make_scene(name, **over)
    p = synth.cluster_params(lib, sc, grid=grid, cap=0)
    miss, members = misses(sc, p)
    assert members > 1000
    assert miss == 0, f"{miss} of {members} members outside their candidate box"


def view_to_world(sc, vpos):
    view = np.array(sc.camera.view, np.float64).reshape(4, 4).T  # row-major
    inv = np.linalg.inv(view)
    h = np.c_[vpos, np.ones(len(vpos))] @ inv.T
    return h[:, :3]


def kernel_view_z(sc, wpos):
    """The view-space z of world positions as the cull computes it (FP32,
    columns accumulated in order)."""
    m = np.array(sc.camera.view, f32).reshape(4, 4)
    x, y, z = (wpos[:, i].astype(f32) for i in range(3))
    return ((m[0, 2] * x + m[1, 2] * y) + m[2, 2] * z) + m[3, 2]


def boundary_scene(lib, seed=5):
    """Point lights whose spheres touch a slice limit: the light's view
    distance is slice_far[k] + rc or slice_near[k] - rc to within a few ulp
    (world positions nudged ulp by ulp until the FP32 view z lands there),
    scattered over the screen."""
    sc = synth.make_scene("c2_1080p_1k", width=320, height=180, num_lights=3000)
    p = synth.cluster_params(lib, sc, grid=(10, 6, 24), cap=0)
    rng = np.random.default_rng(seed)
    n = len(sc.lights)
    L = sc.lights
    L["light_type"] = abi.HL_LIGHT_POINT
    r = rng.uniform(0.2, 3.0, n).astype(f32)
    k = rng.integers(1, p.grid_z - 1, n)
    far_side = rng.random(n) < 0.5
    pad = np.array(p.slice_r_pad[:p.grid_z], f32)[k]
    rc = (r + pad).astype(np.float64)
    sn = np.array(p.slice_near[:p.grid_z], np.float64)[k]
    sf = np.array(p.slice_far[:p.grid_z], np.float64)[k]
    zc = np.where(far_side, sf + rc, sn - rc)
    zc = np.maximum(zc, 0.5)
    p00, p11 = float(sc.camera.proj[0]), float(sc.camera.proj[5])
    ndc = rng.uniform(-0.95, 0.95, (n, 2))
    vpos = np.c_[ndc[:, 0] * zc / p00, ndc[:, 1] * zc / p11, -zc]
    w = view_to_world(sc, vpos).astype(f32)
    # nudge along the view axis, ulp by ulp, towards the target view z
    fwd = np.array(sc.camera.view, np.float64).reshape(4, 4).T[2, :3]  # view z row
    axis = int(np.argmax(np.abs(fwd)))
    for _ in range(64):
        z = kernel_view_z(sc, w).astype(np.float64)
        err = -z - zc
        done = np.abs(err) <= 2e-7 * np.abs(zc)
        if done.all():
            break
        step = np.where(err * fwd[axis] > 0, np.inf, -np.inf).astype(f32)  # move z towards the target
        w[~done, axis] = np.nextafter(w[~done, axis], step[~done])
    L["position_range"][:, :3] = w
    L["position_range"][:, 3] = r
    L["color_intensity"] = (1.0, 1.0, 1.0, 1.0)
    return sc, p


def test_candidates_at_slice_limits(lib):
    sc, p = boundary_scene(lib)
    miss, members = misses(sc, p)
    assert members > 3000
    assert miss == 0, f"{miss} of {members} members outside their candidate box"


@pytest.mark.parametrize("mistake", ["no_pad", "no_tilt", "pad_of_first"])
def test_planted_box_mistakes_are_detected(lib, mistake):
    """A box built without the slice padding, with the lateral planes taken
    as parallel to the view axis, or under the padding of the nearest slice in
    reach misses members of the slice-limit scene.  (Dropping the rounding
    margins misses none here: they cover a few ulp, which no sample crosses;
    they stay for the kernel's fast reciprocals.)"""
    sc, p = boundary_scene(lib)
    miss, members = misses(sc, p, mistake)
    assert miss > 0, f"{mistake}: no member missed"


def test_cluster_params_init_tables_are_sorted(lib):
    """hl_cluster_params_init returns the non-decreasing slice tables that the
    cull requires (and validates) for every scene and grid used here."""
    for name, grid in (("c1_720p_64", None), ("c2_1080p_1k", (16, 9, 24)), ("c3_4k_16k", (32, 18, 48)),
                       ("c3_4k_16k_far", (32, 18, 64)), ("c4_512_adj", (16, 16, 64))):
        sc = synth.make_scene(name, width=320, height=180)
        p = synth.cluster_params(lib, sc, grid=grid, cap=0)
        for k in range(1, p.grid_z):
            assert p.slice_near[k] >= p.slice_near[k - 1] and p.slice_far[k] >= p.slice_far[k - 1]
            assert p.slice_r_pad[k] >= p.slice_r_pad[k - 1] >= 0.0
        assert p.r_pad_column == max(p.slice_r_pad[:p.grid_z])
