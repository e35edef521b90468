"""Exact membership of the clustered lists, pinned from both sides.

The clustered cull has no reference counterpart; its rule (stated with
hl_cluster_params in helio_hip.h and restated in oracle.c or_cluster_member)
is: a light is in cell (column, slice k) iff
  1. its sphere, radius r + pad_k, passes the column's four lateral planes
     (the tile cull's plane construction, light_cull.wgsl:85-117, on the
     column's pixel rectangle);
  2. that sphere meets the cell's view-space AABB (the column's NDC corners
     unprojected at slice_near[k] and slice_far[k]);
  3. a spot light's cone (outer angle) meets the AABB's bounding sphere grown
     by pad_k;
directional lights are in every cell of an on-screen column.

test_cluster_lists_are_conservative (test_oracle.py) pins the lists from
below.  Here an independent FP64 evaluation of the same rule, written from the
statement above, must give the same member set as the oracle for every cell of
small scenes (spot lights included), except for pairs within 1e-5 relative of
one of the rule's boundaries, where FP32 rounding may decide either way.  An
over-inclusive rule (a cone test that always passes, padding ten times too
large) or an under-inclusive one fails.
"""
import numpy as np
import pytest

import oracle_lib as O
from helio_amd import abi, synth


def fp64_members(sc, p):
    """[cells, lights] membership and margin masks of the rule, in float64."""
    cam = sc.camera
    p00, p11 = float(cam.proj[0]), float(cam.proj[5])
    view = np.array(cam.view[:], np.float64).reshape(4, 4).T  # row-major
    L = sc.lights
    n = len(L)
    pos = np.c_[L["position_range"][:, :3].astype(np.float64), np.ones(n)]
    c = (view @ pos.T).T[:, :3]
    a = (view[:3, :3] @ L["direction_outer"][:, :3].astype(np.float64).T).T
    r = L["position_range"][:, 3].astype(np.float64)
    outer = L["direction_outer"][:, 3].astype(np.float64)
    typ = L["light_type"]
    gx, gy, gz = p.grid_x, p.grid_y, p.grid_z
    W, H = p.width, p.height
    ncell = gx * gy * gz
    member = np.zeros((ncell, n), bool)
    margin = np.zeros((ncell, n), bool)
    clen = np.linalg.norm(c, axis=1)

    def plane(q0, q1):
        r0 = np.array([q0[0] / p00, q0[1] / p11, -1.0])
        r1 = np.array([q1[0] / p00, q1[1] / p11, -1.0])
        r0 /= np.linalg.norm(r0)
        r1 /= np.linalg.norm(r1)
        v = np.cross(r0, r1)
        return v / np.linalg.norm(v)

    alen = np.linalg.norm(a, axis=1)
    with np.errstate(all="ignore"):
        cos_t = np.where(alen > 0, outer / alen, 0.0)
    cone = (typ == abi.HL_LIGHT_SPOT) & (alen > 0) & (cos_t > 0)
    cos_t = np.minimum(cos_t, 1.0)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    an = a / np.where(alen > 0, alen, 1.0)[:, None]
    for cy in range(gy):
        for cx in range(gx):
            px0, py0 = cx * p.cell_w, cy * p.cell_h
            if px0 >= W or py0 >= H:
                continue
            px1, py1 = min(px0 + p.cell_w, W), min(py0 + p.cell_h, H)
            l, rr = px0 / W * 2 - 1, px1 / W * 2 - 1
            t, b = 1 - py0 / H * 2, 1 - py1 / H * 2
            P = np.stack([plane((l, b), (l, t)), plane((rr, t), (rr, b)), plane((l, t), (rr, t)),
                          plane((rr, b), (l, b))])
            dpl = c @ P.T  # [n, 4]
            for k in range(gz):
                cell = cx + gx * (cy + gy * k)
                pad = float(p.slice_r_pad[k])
                rc = r + pad
                zn, zf = float(p.slice_near[k]), float(p.slice_far[k])
                xs = np.array([l * zn, rr * zn, l * zf, rr * zf]) / p00
                ys = np.array([b * zn, t * zn, b * zf, t * zf]) / p11
                lo = np.array([xs.min(), ys.min(), -zf])
                hi = np.array([xs.max(), ys.max(), -zn])
                tol = 1e-5 * (clen + rc)
                in_pl = np.all(dpl >= -rc[:, None], axis=1)
                m_pl = np.any(np.abs(dpl + rc[:, None]) < tol[:, None], axis=1)
                dd = np.maximum(np.maximum(lo - c, c - hi), 0.0)
                dist = np.linalg.norm(dd, axis=1)
                in_bx = dist <= rc
                m_bx = np.abs(dist - rc) < tol
                mid, ext = 0.5 * (lo + hi), hi - lo
                R = 0.5 * np.linalg.norm(ext) + pad
                v = mid - c
                v1 = (v * an).sum(1)
                perp = np.sqrt(np.maximum(0.0, (v * v).sum(1) - v1 * v1))
                dcp = cos_t * perp - v1 * sin_t
                ctol = 1e-5 * (np.linalg.norm(v, axis=1) + R)
                in_cn = ~cone | ((dcp <= R) & (v1 >= -R))
                m_cn = cone & ((np.abs(dcp - R) < ctol) | (np.abs(v1 + R) < ctol))
                directional = typ == abi.HL_LIGHT_DIRECTIONAL
                member[cell] = directional | (in_pl & in_bx & in_cn)
                margin[cell] = ~directional & (m_pl | m_bx | m_cn)
    return member, margin


def oracle_members(sc, p):
    off, cnt, pool, flags = O.cull_clusters(sc, p, 1 << 22)
    assert flags == 0
    m = np.zeros((len(cnt), len(sc.lights)), bool)
    for cell in range(len(cnt)):
        m[cell, pool[off[cell]:off[cell] + cnt[cell]]] = True
    return m


@pytest.mark.parametrize("name,over,grid", [
    ("c2_1080p_1k", dict(width=160, height=96, num_lights=300), (8, 6, 12)),
    ("c2_1080p_1k", dict(width=128, height=80, num_lights=200, seed=synth.SEED_BASE + 901), (5, 4, 24)),
    ("c1_720p_64", dict(width=96, height=64), (4, 3, 16)),
])
def test_membership_matches_fp64_rule(hlib, name, over, grid):
    sc = synth.make_scene(name, **over)
    p = synth.cluster_params(hlib, sc, grid=grid, cap=0)
    got = oracle_members(sc, p)
    want, margin = fp64_members(sc, p)
    diff = got ^ want
    assert want.sum() > 500 and (sc.lights["light_type"] == abi.HL_LIGHT_SPOT).any() or name == "c1_720p_64"
    assert not (diff & ~margin).any(), (
        f"{int((diff & ~margin).sum())} pairs differ outside the margins "
        f"(oracle extra {int((got & ~want & ~margin).sum())}, missing {int((want & ~got & ~margin).sum())})")
    assert diff.sum() <= 0.001 * want.sum() + 2


def test_fp64_rule_rejects_cone_and_padding_mistakes(hlib):
    """The FP64 evaluation above is sensitive to the rule's parts: evaluated
    without the cone test, or with the padding ten times larger, it no longer
    matches the oracle outside the margins."""
    sc = synth.make_scene("c2_1080p_1k", width=160, height=96, num_lights=300)
    p = synth.cluster_params(hlib, sc, grid=(8, 6, 12), cap=0)
    got = oracle_members(sc, p)
    lights = sc.lights.copy()
    sc_nc = synth.make_scene("c2_1080p_1k", width=160, height=96, num_lights=300)
    lights["light_type"][lights["light_type"] == abi.HL_LIGHT_SPOT] = abi.HL_LIGHT_POINT
    sc_nc.lights = lights
    want, margin = fp64_members(sc_nc, p)
    assert ((got ^ want) & ~margin).any(), "dropping the cone test must be detected"
    p10 = synth.cluster_params(hlib, sc, grid=(8, 6, 12), cap=0)
    for k in range(p10.grid_z):
        p10.slice_r_pad[k] = 10.0 * p10.slice_r_pad[k] + 0.05
    want, margin = fp64_members(sc, p10)
    assert ((got ^ want) & ~margin).any(), "a ten-fold padding must be detected"
