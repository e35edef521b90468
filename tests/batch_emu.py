"""CPU emulation of the index arithmetic of a batched multi-view frame
(helio_amd/csrc/hl_batch.hip build_plan + enqueue_frame, the k_batch_* cull
kernels of hl_cull.hip and the k_shade_batch item walk of hl_shade.hip).

Every address a kernel forms is recomputed here from the same formulas, with
the real cameras, light subsets and cluster grids of a plan, and asserted
against the size of the buffer the plan allocates for it.  Floating-point
decisions that only select *which* in-range element is touched (membership,
the slice guess) are not needed for that and are replaced by their
worst cases (every candidate cell a member, any raw count).

Two cell layouts are covered:
  aligned=True   each view's cells start on a multiple of kBatchCellAlign
                 (256): a 256-thread block of the geometry kernel and a
                 16-cluster block of count/fill lie in one view;
  aligned=False  views packed back to back (the layout of the batched cull
                 before the alignment): count/fill look the view up per
                 16-lane group, and a block may straddle two views.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from helio_amd import abi, synth

# constants of hl_internal.hpp / hl_cull.hip / hl_shade.hip
K_MARK_INLINE = 2
K_MARK_QUEUES = 64
K_MARK_COUNT_STRIDE = 32
K_SCAN_TILE = 1024
K_CELL_LANES = 16
K_BATCH_CELL_ALIGN = 256
K_SHADE_QUEUES = 64
K_GRAD_SLOTS = 8
K_FUSED_SPLIT_BLOCKS = 4096
SIZEOF_COLUMN_GEOM = 80
SIZEOF_CELL_BOX = 24
SIZEOF_HOT = 80
SIZEOF_CULL = 64


def mark_queue_seg(num_lights: int) -> int:
    return (4 * num_lights + 4096 + K_MARK_QUEUES - 1) // K_MARK_QUEUES


def geom_offset_boxes(ncol: int) -> int:
    return (SIZEOF_COLUMN_GEOM * ncol + 255) & ~255


def tile_sums_offset(ncell: int) -> int:
    return (4 * ncell + 255) & ~255


def scan_tiles(ncell: int) -> int:
    return (ncell + K_SCAN_TILE - 1) // K_SCAN_TILE


@dataclass
class EmuView:
    width: int
    height: int
    camera: abi.GpuCameraUniforms
    lights: np.ndarray          # GPU_LIGHT_DTYPE records of the view's subset (pool order)
    subset: np.ndarray          # ascending pool indices
    adjoint: bool
    params: abi.hl_cluster_params | None = None

    @property
    def n(self) -> int:
        return len(self.subset)


def find_view(base: np.ndarray, g: np.ndarray) -> np.ndarray:
    """hl_cull.hip find_view: the view v with base[v] <= g < base[v + 1]
    (the last such v when views are empty), vectorised."""
    return np.searchsorted(base[1:-1], g, side="right").astype(np.int64)


def light_boxes(v: EmuView, mistake: str | None = None) -> tuple[np.ndarray, ...]:
    """light_cells() of hl_cull.hip for every light of the view: the candidate
    box (x0, y0, s0, nx, ny, nz) and whether it is non-empty.  The fast
    arithmetic is evaluated in float32 numpy; the integer clamps that follow
    it are the kernel's, and they alone keep the box inside the grid.
    mistake plants an error for tests: "no_margins" (no rounding margins),
    "no_pad" (the radius without the slice padding), "no_tilt" (fx = fy = 1,
    the lateral planes taken as parallel to the view axis), "pad_of_first"
    (the padding of the nearest slice in reach instead of the farthest)."""
    margins = mistake != "no_margins"
    p = v.params
    gx, gy, gz = p.grid_x, p.grid_y, p.grid_z
    cam = v.camera
    view = np.array(cam.view, np.float32).reshape(4, 4)  # column-major: view[col][row]
    pos = np.stack([v.lights["position_range"][:, k] for k in range(3)], 1).astype(np.float32)
    rng = np.where(v.lights["light_type"] == 0, np.float32(np.inf),
                   v.lights["position_range"][:, 3]).astype(np.float32)
    # view-space position (x, y, z), columns accumulated in order (hl_x_mat_row)
    vp = np.stack([((view[0, r] * pos[:, 0] + view[1, r] * pos[:, 1]) + view[2, r] * pos[:, 2]) + view[3, r]
                   for r in range(3)], 1).astype(np.float32)
    sn = np.array(p.slice_near[:gz], np.float32)
    sf = np.array(p.slice_far[:gz], np.float32)
    sp = np.array(p.slice_r_pad[:gz], np.float32)
    n = len(rng)
    x0 = np.zeros(n, np.int64)
    x1 = np.full(n, gx - 1, np.int64)
    y0 = np.zeros(n, np.int64)
    y1 = np.full(n, gy - 1, np.int64)
    s0 = np.zeros(n, np.int64)
    s1 = np.full(n, gz - 1, np.int64)
    zc = (-vp[:, 2]).astype(np.float32)
    f32 = np.float32
    with np.errstate(all="ignore"):
        fin0 = ~np.isinf(rng) & np.isfinite(zc) & np.isfinite(vp[:, 0]) & np.isfinite(vp[:, 1])
        fin = fin0.copy()
        e = f32(1e-5) if margins else f32(0)
        snm = (sn - e * np.abs(sn)).astype(f32)
        sfm = (sf + e * np.abs(sf)).astype(f32)
        slack = (e * ((np.abs(vp[:, 0]) + np.abs(vp[:, 1])) + np.abs(zc))).astype(f32)
        r = np.abs(rng).astype(f32)
        spk = sp * f32(0) if mistake == "no_pad" else sp
        rr = ((r + spk[gz - 1]) * (f32(1) + e) + slack).astype(f32)
        t1 = (snm[None, :] <= (zc + rr)[:, None]).sum(1) - 1
        fin &= t1 >= 0
        rr = ((r + spk[np.maximum(t1, 0)]) * (f32(1) + e) + slack).astype(f32)
        if mistake == "pad_of_first":
            f0 = np.minimum((sfm[None, :] < (zc - rr)[:, None]).sum(1), gz - 1)
            rr = ((r + spk[f0]) * (f32(1) + e) + slack).astype(f32)
        ls0 = (sfm[None, :] < (zc - rr)[:, None]).sum(1)
        ls1 = (snm[None, :] <= (zc + rr)[:, None]).sum(1) - 1
        s0 = np.where(fin, ls0, s0)
        s1 = np.where(fin, ls1, s1)
        none = fin0 & (t1 < 0)  # no slice in reach: no candidates
        p00, p11 = f32(cam.proj[0]), f32(cam.proj[5])
        front = fin & (zc > f32(1e-4)) & (p00 > 0) & (p11 > 0)
        fx = np.sqrt(f32(1) + f32(1) / (p00 * p00))
        fy = np.sqrt(f32(1) + f32(1) / (p11 * p11))
        if mistake == "no_tilt":
            fx = fy = f32(1)
        iz = f32(1) / zc
        nx0 = p00 * (vp[:, 0] - rr * fx) * iz
        nx1 = p00 * (vp[:, 0] + rr * fx) * iz
        ny0 = p11 * (vp[:, 1] - rr * fy) * iz
        ny1 = p11 * (vp[:, 1] + rr * fy) * iz
        fw, fh = f32(p.width), f32(p.height)
        px0 = (nx0 + 1) * f32(0.5) * fw
        px1 = (nx1 + 1) * f32(0.5) * fw
        py0 = (1 - ny1) * f32(0.5) * fh
        py1 = (1 - ny0) * f32(0.5) * fh
        q = f32(0.25) if margins else f32(0)
        px0 = px0 - (e * np.abs(px0) + q)
        px1 = px1 + (e * np.abs(px1) + q)
        py0 = py0 - (e * np.abs(py0) + q)
        py1 = py1 + (e * np.abs(py1) + q)
        icw, ich = f32(1) / f32(p.cell_w), f32(1) / f32(p.cell_h)

        def to_int(f):  # v_cvt_i32_f32 of a value already clamped to [-1, g]
            return np.trunc(f.astype(np.float64)).astype(np.int64)

        fmax = np.fmax  # fmaxf/fminf: a NaN operand yields the other one
        fmin = np.fmin
        bx0 = to_int(fmin(fmax(np.floor(px0 * icw), f32(-1)), f32(gx)))
        bx1 = to_int(fmax(fmin(np.floor(px1 * icw), f32(gx)), f32(-1)))
        by0 = to_int(fmin(fmax(np.floor(py0 * ich), f32(-1)), f32(gy)))
        by1 = to_int(fmax(fmin(np.floor(py1 * ich), f32(gy)), f32(-1)))
    x0 = np.where(front, bx0, x0)
    x1 = np.where(front, bx1, x1)
    y0 = np.where(front, by0, y0)
    y1 = np.where(front, by1, y1)
    x0 = np.clip(x0, 0, gx - 1)
    x1 = np.clip(x1, -1, gx - 1)
    y0 = np.clip(y0, 0, gy - 1)
    y1 = np.clip(y1, -1, gy - 1)
    s0 = np.clip(s0, 0, gz - 1)
    s1 = np.clip(s1, -1, gz - 1)
    ok = (x1 >= x0) & (y1 >= y0) & (s1 >= s0) & ~none
    return ok, x0, y0, s0, x1 - x0 + 1, y1 - y0 + 1, s1 - s0 + 1


def floor_div_f32(x: np.ndarray, c: int) -> np.ndarray:
    """hl_floor_div(x, c) = (uint)fma(x, r, 0.5 r) for both reciprocals r one
    ulp either side of the rounded 1/c (v_rcp_f32 is within 1 ulp); returns
    the two results."""
    r = np.float32(1.0) / np.float32(c)
    out = []
    for rr in (np.nextafter(r, np.float32(0)), np.nextafter(r, np.float32(1))):
        xf = x.astype(np.float64)
        q = xf * np.float64(rr) + 0.5 * np.float64(rr)  # exact product-sum, one rounding below
        out.append(np.floor(q.astype(np.float32)).astype(np.int64))
    return out


def build_plan(views: list[EmuView], cfgs: list[tuple[int, int, int]], lib, aligned: bool = True) -> dict:
    """build_plan's host sizing for views with configurations (tw_log2, cell,
    slices); fills each view's cluster params.  Returns the plan's bases,
    sizes and per-view allocations (bytes)."""
    nv = len(views)
    light_base = np.zeros(nv + 1, np.int64)
    cell_base = np.zeros(nv + 1, np.int64)
    tile_base = np.zeros(nv + 1, np.int64)
    alloc = []
    for i, (v, (tw, cell, gz)) in enumerate(zip(views, cfgs)):
        gx, gy = -(-v.width // cell), -(-v.height // cell)
        p = abi.hl_cluster_params()
        lib.call("hl_cluster_params_init", C.byref(v.camera), v.width, v.height, gx, gy, gz, 0, C.byref(p))
        v.params = p
        ncol, ncell = gx * gy, gx * gy * gz
        assert ncol <= 65535 and ncell <= 1 << 26, "check_cluster_params bounds"
        n = v.n
        W = (n + 31) // 32
        SW = (W + 31) // 32
        pool_cap = min(ncell * min(n, 32) + 4096, 0x7FFFFFFF)
        alloc.append(dict(ncol=ncol, ncell=ncell, W=W, SW=SW, tw=tw, hot=SIZEOF_HOT * max(n, 1),
                          cull=SIZEOF_CULL * max(n, 1), geom=geom_offset_boxes(ncol) + SIZEOF_CELL_BOX * ncell,
                          bits=4 * ncell * W, summary=4 * ncell * SW,
                          raw=tile_sums_offset(ncell) + 8 * scan_tiles(ncell),
                          offsets=4 * ncell, counts=4 * ncell, pool_cap=pool_cap,
                          acc=8 * K_GRAD_SLOTS * 4 * max(n, 1) if v.adjoint else 0))
        light_base[i + 1] = light_base[i] + n
        span = -(-ncell // K_BATCH_CELL_ALIGN) * K_BATCH_CELL_ALIGN if aligned else ncell
        cell_base[i + 1] = cell_base[i] + span
        tile_base[i + 1] = tile_base[i] + scan_tiles(ncell)
    assert light_base[-1] < 1 << 28 and cell_base[-1] < 1 << 31
    seg = mark_queue_seg(int(light_base[-1]))
    markq_bytes = 4 * K_MARK_QUEUES * K_MARK_COUNT_STRIDE + 16 * int(light_base[-1]) + 8 * K_MARK_QUEUES * seg
    return dict(light_base=light_base, cell_base=cell_base, tile_base=tile_base, alloc=alloc, seg=seg,
                markq_bytes=markq_bytes)


def check_geom(views, plan, aligned=True) -> None:
    """k_batch_geom: the cell blocks (every thread's view, cell and the
    records it writes) and the light-gather blocks behind them."""
    cb, lb, al = plan["cell_base"], plan["light_base"], plan["alloc"]
    total_cells = int(cb[-1])
    if aligned:
        assert total_cells % K_BATCH_CELL_ALIGN == 0
        geom_blocks = total_cells // 256
        g = np.arange(geom_blocks * 256, dtype=np.int64)
        vi = find_view(cb, (g // 256) * 256)  # one lookup per block
    else:
        geom_blocks = -(-total_cells // 256)
        g = np.arange(geom_blocks * 256, dtype=np.int64)
        g = g[g < total_cells]
        vi = find_view(cb, g)
    c = g - cb[vi]
    ncell = np.array([a["ncell"] for a in al])[vi]
    ncol = np.array([a["ncol"] for a in al])[vi]
    assert (c >= 0).all()
    live = c < ncell
    # tile sums: c < scan_tiles(ncell) entries of 8 bytes after the raw counts
    ts = c < (ncell + K_SCAN_TILE - 1) // K_SCAN_TILE
    raw_bytes = np.array([a["raw"] for a in al])[vi]
    ts_off = np.array([tile_sums_offset(a["ncell"]) for a in al])[vi]
    assert (ts_off[ts] + 8 * (c[ts] + 1) <= raw_bytes[ts]).all()
    # boxes[c] for live cells, geom[col] for slice 0
    geom_bytes = np.array([a["geom"] for a in al])[vi]
    box_off = np.array([geom_offset_boxes(a["ncol"]) for a in al])[vi]
    assert (box_off[live] + SIZEOF_CELL_BOX * (c[live] + 1) <= geom_bytes[live]).all()
    col = c % ncol
    k0 = live & (c < ncol)
    assert (SIZEOF_COLUMN_GEOM * (col[k0] + 1) <= box_off[k0]).all()
    if aligned:  # every block lies in one view
        blk = g // 256
        first = vi[blk * 256 == g]
        assert np.array_equal(np.repeat(first, 256), vi)
    # light gather: l < total lights, subset entry and records in range
    ltot = int(lb[-1])
    l = np.arange(ltot, dtype=np.int64)
    lv = find_view(lb, l)
    li = l - lb[lv]
    nvl = np.array([v.n for v in views])[lv]
    assert ((li >= 0) & (li < nvl)).all()


_FLOOR_DIV_CHECKED: dict[int, int] = {}


def assert_floor_div_exact(divisor: int, limit: int) -> None:
    """hl_floor_div(x, divisor) == x // divisor for every x < limit (limit <=
    2^21), with the reciprocal one ulp either side; cached per divisor."""
    assert limit <= 1 << 21
    if _FLOOR_DIV_CHECKED.get(divisor, 0) >= limit:
        return
    x = np.arange(limit, dtype=np.int64)
    for q in floor_div_f32(x, divisor):
        assert np.array_equal(q, x // divisor), divisor
    _FLOOR_DIV_CHECKED[divisor] = limit


def check_mark(views, plan) -> dict:
    """k_batch_mark + k_batch_mark_tail: each light's candidate box, its
    queue claims (the bounds hold in any claim order), the packed entries,
    and the cell, bitmask and summary index of every candidate it tests.
    A candidate t of a box decomposes as t = (iy * nx + ix) * nz + kz by two
    hl_floor_div steps (exact below 2^21, integer division above): with the
    box inside the grid, every cell index lies below the view's cell count."""
    lb, al = plan["light_base"], plan["alloc"]
    seg = plan["seg"]
    counts = np.zeros(K_MARK_QUEUES, np.int64)
    entries = [[] for _ in range(K_MARK_QUEUES)]
    bounds = {}
    stats = dict(lights=0, chunks=0, queued=0, max_work=0)
    for vi, (v, a) in enumerate(zip(views, al)):
        p = v.params
        ok, x0, y0, s0, nx, ny, nz = light_boxes(v)
        assert (x0 >= 0).all() and (y0 >= 0).all() and (s0 >= 0).all()
        assert ((x0 + nx <= p.grid_x) | ~ok).all() and ((y0 + ny <= p.grid_y) | ~ok).all()
        assert ((s0 + nz <= p.grid_z) | ~ok).all()
        assert (nx < 1 << 16).all() and (ny < 1 << 16).all() and (nz < 1 << 16).all() and (s0 < 1 << 16).all()
        work = nx * ny * nz
        nchunk = (work + 63) // 64
        assert (nchunk[ok] < 1 << 24).all()
        if ok.any():
            stats["max_work"] = max(stats["max_work"], int(work[ok].max()))
            # the highest cell any candidate can reach, and its bitmask words
            cmax = (x0 + nx - 1) + p.grid_x * (y0 + ny - 1) + p.grid_x * p.grid_y * (s0 + nz - 1)
            assert (cmax[ok] < a["ncell"]).all()
            li = np.nonzero(ok)[0]
            # word-major bitmask: word w of cluster c at w * ncell + c
            assert (4 * ((li >> 5) * a["ncell"] + cmax[ok] + 1) <= a["bits"]).all()
            assert (4 * ((li >> 10) * a["ncell"] + cmax[ok] + 1) <= a["summary"]).all()
            small = ok & (work <= 1 << 21)
            for d in np.unique(nz[small]):
                assert_floor_div_exact(int(d), int(work[small & (nz == d)].max()))
            for d in np.unique(nx[small]):
                assert_floor_div_exact(int(d), int((work // np.maximum(nz, 1))[small & (nx == d)].max()))
        stats["lights"] += int(ok.sum())
        stats["chunks"] += int(nchunk[ok].sum())
        for li in np.nonzero(ok & (nchunk > K_MARK_INLINE))[0]:
            g = int(lb[vi] + li)
            extra = int(nchunk[li]) - K_MARK_INLINE
            sub = g % K_MARK_QUEUES
            base = int(counts[sub])
            counts[sub] += extra
            fit = 0 if base >= seg else min(extra, seg - base)
            if fit:
                bounds[g] = (int(nx[li] * ny[li] * nz[li]),)
                assert base + fit <= seg
                entries[sub].extend((g, (vi << 24) | (K_MARK_INLINE + j)) for j in range(fit))
                stats["queued"] += fit
    # the queue's memory: counters, bounds[total lights], entries[64 * seg];
    # the tail reads entries [0, min(count, seg)), all written this frame
    assert plan["markq_bytes"] >= 4 * K_MARK_QUEUES * K_MARK_COUNT_STRIDE + 16 * int(lb[-1]) + \
        8 * K_MARK_QUEUES * seg
    for sub in range(K_MARK_QUEUES):
        assert len(entries[sub]) == min(int(counts[sub]), seg)
        for g, packed in entries[sub]:
            vi, chunk = packed >> 24, packed & 0xFFFFFF
            assert vi < len(views)
            li = g - int(lb[vi])
            assert 0 <= li < views[vi].n and g in bounds and g < int(lb[-1])
            assert chunk < -(-bounds[g][0] // 64)
    stats["queue_max"] = int(counts.max())
    return stats


def scan_emulate(raw: np.ndarray, cap: int, pool_cap: int) -> tuple[np.ndarray, np.ndarray]:
    """scan_tile over every tile of a view: offsets / counts as the kernel
    writes them (capped, pool-truncated)."""
    v = raw.astype(np.int64)
    if cap:
        v = np.minimum(v, cap)
    off = np.concatenate([[0], np.cumsum(v)[:-1]])
    o = np.minimum(off, pool_cap)
    room = pool_cap - o
    cnt = np.minimum(v, room)
    return o, cnt


def check_count_scan_fill(views, plan, aligned=True, rng=None) -> None:
    """k_batch_count / k_batch_fill: each 16-lane group's view and cluster,
    the bitmask words it may list (word indices below W: a summary bit is set
    only for a word holding a member), the tile sums and raw counts the fill
    block reads for its offsets, and the pool positions the fill writes for
    any raw counts."""
    cb, tb, al = plan["cell_base"], plan["tile_base"], plan["alloc"]
    total_cells = int(cb[-1])
    cblocks = total_cells // 16 if aligned else -(-total_cells // 16)
    grp = np.arange(cblocks * 16, dtype=np.int64)
    if aligned:
        vi = find_view(cb, (grp // 16) * 16)
    else:
        on = grp < total_cells
        vi = find_view(cb, np.where(on, grp, total_cells - 1))
    c = grp - cb[vi]
    ncell = np.array([a["ncell"] for a in al])[vi]
    if not aligned:
        ncell = np.where(grp < total_cells, ncell, 0)
    live = (c >= 0) & (c < ncell)
    W = np.array([a["W"] for a in al])[vi]
    SW = np.array([a["SW"] for a in al])[vi]
    bits = np.array([a["bits"] for a in al])[vi]
    summ = np.array([a["summary"] for a in al])[vi]
    # summary words j < SW of the cluster, listed word indices 32 j + b < W
    # (word-major: word w of cluster c at w * ncell + c)
    assert (4 * ((SW[live] - 1) * ncell[live] + c[live] + 1) <= summ[live]).all()
    assert (4 * ((W[live] - 1) * ncell[live] + c[live] + 1) <= bits[live]).all()
    # a group with no cluster reads only summary word range of cluster 0 with on=false (no loads)
    # scan inside the fill (scan_fill_block): a block whose first cluster f
    # lies in the view reads the sums of scan tiles [0, f // 1024) and the raw
    # counts [1024 (f // 1024), f) with 16-B loads; padding blocks return.
    # (Views packed back to back, the layout before the alignment, had a
    # separate scan launch: nothing to check there.)
    if not aligned:
        return check_pool_positions(views, al, rng)
    first = (np.arange(cblocks, dtype=np.int64) * 16)
    bv = vi[::16]
    f = first - cb[bv]
    nc = np.array([a["ncell"] for a in al])[bv]
    live_b = (f >= 0) & (f < nc)
    assert (f[live_b] % 16 == 0).all()
    ntile = np.array([scan_tiles(a["ncell"]) for a in al])[bv]
    assert ((f[live_b] // 1024) <= ntile[live_b]).all()
    raw_b = np.array([a["raw"] for a in al])[bv]
    ts_off = np.array([tile_sums_offset(a["ncell"]) for a in al])[bv]
    # the last tile sum read, and the last raw word read (f - 1 < ncell), inside raw
    assert ((ts_off + 8 * (f // 1024))[live_b] <= raw_b[live_b]).all()
    assert (4 * f[live_b] <= ts_off[live_b]).all()
    check_pool_positions(views, al, rng)


def check_pool_positions(views, al, rng=None) -> None:
    """The pool positions the fill writes, for extreme and random raw counts."""
    rng = rng or np.random.default_rng(1)
    for v, a in zip(views, al):
        n = v.n
        for raw in (np.full(a["ncell"], n), rng.integers(0, n + 1, a["ncell"]),
                    np.zeros(a["ncell"], np.int64)):
            o, cnt = scan_emulate(raw, v.params.cap, a["pool_cap"])
            assert (o + cnt <= a["pool_cap"]).all()


def check_shade_items(views, plan, cus: int = 256) -> None:
    """The persistent shading launches: per (format set, mode) group the
    items of its views (heaviest first), each item's span by bisection, the
    block tile's pixels and the records it touches."""
    al = plan["alloc"]
    import math
    order = sorted(range(len(views)), key=lambda i: -((4.0 if views[i].adjoint else 1.0) *
                                                      (1.0 + math.log2(1.0 + views[i].n))))
    for adj in (False, True):
        spans = []
        nitems = 0
        for i in order:
            if views[i].adjoint != adj:
                continue
            tw = al[i]["tw"]
            bw, bh = 4 << tw, 64 >> tw
            # adjoint views below K_FUSED_SPLIT_BLOCKS block tiles: one item per
            # wave tile (shade_items), four per block tile; larger ones one per
            # block tile
            blocks = -(-views[i].width // bw) * -(-views[i].height // bh)
            split = adj and blocks < K_FUSED_SPLIT_BLOCKS
            nb = blocks * (4 if split else 1)
            spans.append((nitems, i, nb, bw, bh, split))
            nitems += nb
        if not spans:
            continue
        assert nitems < 1 << 31
        starts = np.array([s[0] for s in spans] + [nitems])
        it = np.arange(nitems, dtype=np.int64)
        lo = np.searchsorted(starts[1:-1], it, side="right")
        for k, (first, i, nb, bw, bh, split) in enumerate(spans):
            m = lo == k
            b = it[m] - first
            assert (b >= 0).all() and (b < nb).all()
            v = views[i]
            btx = -(-v.width // bw)
            bt = b >> 2 if split else b
            bx, by = bt % btx, bt // btx
            x_max = bx.max() * bw + bw - 1
            y_max = by.max() * bh + bh - 1
            # the tiles cover the view; pixels past the edge are masked (inside test)
            assert x_max >= v.width - 1 and y_max >= v.height - 1
            assert x_max < v.width + bw and y_max < v.height + bh
            if v.adjoint:  # gradient slot index of any light: slot * n + li < 8 n
                slot = (b * 4 + 3) % K_GRAD_SLOTS
                assert 8 * (4 * (int(slot.max()) * v.n + v.n - 1) + 4) <= al[i]["acc"]


def c5_views(lib, small: bool = False) -> tuple[list[EmuView], int]:
    specs, subsets, pool_spec = synth.multiview_specs(small=small)
    pool = synth.make_lights(pool_spec)
    views = []
    for s, sub in zip(specs, subsets):
        views.append(EmuView(width=s.width, height=s.height, camera=synth.make_camera(s), lights=pool[sub],
                             subset=sub, adjoint=s.name.startswith("eye")))
    return views, len(pool)


def candidates() -> list[tuple[int, int, int]]:
    """The autotuner's candidates (hl_batch.hip candidates(false)), default first."""
    out = [(3, 64, 24)]
    for tw in (3, 4):
        for cell in (32, 64, 128):
            for gz in (16, 24, 32, 48):
                if (tw, cell, gz) != (3, 64, 24):
                    out.append((tw, cell, gz))
    return out
