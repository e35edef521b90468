"""Golden fixtures: the oracle reproduces them bit for bit (CPU); the HIP path
matches them (GPU: lists bit-exact, radiance within 1e-4 rel + 1e-6)."""
import ctypes as C

import numpy as np
import pytest

import golden_util as G
import oracle_lib as O

NAMES = G.names()


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    sc, params, exp = G.load(name)
    tl, tc = O.cull_tiles(sc)
    assert np.array_equal(tc, exp["tile_counts"])
    for t in range(len(tc)):
        assert np.array_equal(tl[t, :tc[t]], exp["tile_lists"][t, :tc[t]])
    flat = np.ascontiguousarray(tl.reshape(-1))
    out_t, q_t = O.shade(sc, O.tile_grid_host(sc, flat, tc))
    assert q_t == int(exp["q_tiles"])
    assert np.array_equal(out_t.view(np.uint32), exp["radiance_tiles"].view(np.uint32))
    off, cnt, pool, flags = O.cull_clusters(sc, params, 1 << 20)
    assert flags == int(exp["cl_flags"])
    assert np.array_equal(off, exp["cl_offsets"]) and np.array_equal(cnt, exp["cl_counts"])
    assert np.array_equal(pool[:len(exp["cl_pool"])], exp["cl_pool"])
    pool_e = np.ascontiguousarray(exp["cl_pool"])
    out_c, q_c = O.shade(sc, O.cluster_grid_host(params, off, cnt, pool_e))
    assert q_c == int(exp["q_clusters"])
    assert np.array_equal(out_c.view(np.uint32), exp["radiance_clusters"].view(np.uint32))


def _close(hip, ref, mask):
    err = np.abs(hip[mask].astype(np.float64) - ref[mask])
    assert np.all(err <= 1e-4 * np.abs(ref[mask]) + 1e-6), float((err / (np.abs(ref[mask]) + 1e-30)).max())


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_matches_golden(ctx, name):
    from helio_amd.device import DeferredLightPass, LightCullPass, upload_scene
    sc, params, exp = G.load(name)
    print(f"[golden] {name}", flush=True)
    ds = upload_scene(ctx, sc)
    cull = LightCullPass(ctx, sc.width, sc.height)
    cull.prepare(ds)
    cull.execute(ds)
    tc = cull.tiles.counts.download(np.uint32)
    tl = cull.tiles.lists.download(np.uint32).reshape(-1, 64)
    assert np.array_equal(tc, exp["tile_counts"])
    for t in range(len(tc)):
        assert np.array_equal(tl[t, :tc[t]], exp["tile_lists"][t, :tc[t]])
    shade = DeferredLightPass(ctx, sc.width, sc.height)
    shade.prepare(ds)
    shade.out.fill_u32(0)
    shade.execute(ds, cull.publish())
    sky = sc.depth >= 1.0
    _close(shade.output(), exp["radiance_tiles"], ~sky)
    assert shade.count_pairs(ds, cull.publish()) == int(exp["q_tiles"])

    ccull = LightCullPass(ctx, sc.width, sc.height, clusters=params, pool_cap=1 << 20)
    ccull.prepare(ds)
    ccull.execute(ds)
    off, cnt, pool = ccull.cgrid.download()
    assert ccull.overflow() == int(exp["cl_flags"])
    assert np.array_equal(off, exp["cl_offsets"]) and np.array_equal(cnt, exp["cl_counts"])
    assert np.array_equal(pool[:len(exp["cl_pool"])], exp["cl_pool"])
    shade.out.fill_u32(0)
    shade.execute(ds, ccull.publish())
    _close(shade.output(), exp["radiance_clusters"], ~sky)
    assert shade.count_pairs(ds, ccull.publish()) == int(exp["q_clusters"])
