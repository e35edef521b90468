"""Forward-lit fragment shading on the GPU (hl_forward_lit) against the FP32
restatement (oracle/forward_lit.c) on the same light lists: radiance within
1e-4 relative (+1e-6 absolute), alpha and discards exact, pair counts exact."""
import numpy as np
import pytest

from helio_amd import abi, synth
from helio_amd.device import ForwardLitPass, LightCullPass, upload_scene
import oracle_lib as O
from test_gpu_parity import check_radiance

pytestmark = pytest.mark.gpu


class Setup:
    """A scene with SceneDB-style light transforms, its culled lists (on the
    GPU, downloaded for the oracle) and the scene's fragments."""

    def __init__(self, ctx, hlib, name, mode, order="tiles", **over):
        sc = synth.make_scene(name, **over)
        sc.lights, self.entity, self.transforms = synth.make_light_transforms(sc.lights, seed=4)
        self.sc = sc
        self.ds = upload_scene(ctx, sc)
        if mode == "tiles":
            cull = LightCullPass(ctx, sc.width, sc.height)
        else:
            self.params = synth.cluster_params(hlib, sc)
            cull = LightCullPass(ctx, sc.width, sc.height, clusters=self.params, pool_cap=1 << 22)
        cull.prepare(self.ds)
        cull.execute(self.ds)
        self.cull = cull
        self.grid = cull.publish()
        if mode == "tiles":
            lists = cull.tiles.lists.download(np.uint32)
            counts = cull.tiles.counts.download(np.uint32)
            self.host_grid = O.tile_grid_host(sc, lists, counts)
            self.host_grid._arrays = (lists, counts)
        else:
            assert cull.overflow() == 0
            off, cnt, pool = cull.cgrid.download()
            self.host_grid = O.cluster_grid_host(self.params, off, cnt, pool)
        self.fr = synth.make_fragments(sc, order)
        self.fwd = ForwardLitPass(ctx, sc.width, sc.height)
        self.fwd.prepare(self.ds)
        self.ctx = ctx

    def run(self, fr=None, entity=None, use_transforms=True):
        fr = self.fr if fr is None else fr
        ent = (self.entity if entity is None else entity) if use_transforms else None
        d_fr = self.ctx.upload(fr) if len(fr) else self.ctx.alloc(16)
        d_ent = self.ctx.upload(ent) if ent is not None else None
        d_tr = self.ctx.upload(self.transforms) if ent is not None else None
        ntr = len(self.transforms) if ent is not None else 0
        self.fwd.execute(self.ds, d_fr, len(fr), self.grid, entity=d_ent, transforms=d_tr, num_transforms=ntr)
        got = self.fwd.output(len(fr)) if len(fr) else np.zeros((0, 4), np.float32)
        pairs = self.fwd.count_pairs(self.ds, d_fr, len(fr), self.grid, entity=d_ent, transforms=d_tr,
                                     num_transforms=ntr)
        return got, pairs

    def oracle(self, fr=None, entity=None, use_transforms=True):
        fr = self.fr if fr is None else fr
        ent = (self.entity if entity is None else entity) if use_transforms else None
        return O.forward_lit(self.sc.camera, synth.forward_globals(self.sc), fr, self.sc.lights, self.host_grid,
                             ent, self.transforms if ent is not None else None)


def check(got, want):
    check_radiance(got[:, :3], want[:, :3])
    assert np.array_equal(got[:, 3], want[:, 3], equal_nan=True)


@pytest.mark.parametrize("name,mode", [("c1_720p_64", "tiles"), ("c1_720p_64", "clusters"),
                                       ("c2_1080p_1k", "tiles"), ("c2_1080p_1k", "clusters")])
@pytest.mark.parametrize("use_transforms", [True, False])
def test_forward_lit_matches_oracle(ctx, hlib, name, mode, use_transforms):
    print(f"[forward-lit] {name} {mode} transforms={use_transforms}", flush=True)
    s = Setup(ctx, hlib, name, mode)
    got, pairs = s.run(use_transforms=use_transforms)
    want, opairs = s.oracle(use_transforms=use_transforms)
    check(got, want)
    assert pairs == opairs > 0
    assert np.mean(np.any(got[:, :3] > 0, axis=1)) > 0.5


def test_forward_lit_fragment_order_does_not_matter(ctx, hlib):
    """A fragment sums exactly the lights of its own list that pass its range
    test, whatever the other fragments of its wavefront: any order gives the
    same bits."""
    print("[forward-lit] order invariance c2 clusters", flush=True)
    s = Setup(ctx, hlib, "c2_1080p_1k", "clusters")
    got, pairs = s.run()
    perm = np.random.default_rng(9).permutation(len(s.fr))
    got_p, pairs_p = s.run(fr=s.fr[perm])
    assert np.array_equal(got_p, got[perm]) and pairs_p == pairs


def test_forward_lit_discards_and_pixels_outside_the_grid(ctx, hlib):
    print("[forward-lit] discards, pixels outside the grid", flush=True)
    s = Setup(ctx, hlib, "c1_720p_64", "tiles")
    fr = s.fr.copy()
    fr["alpha"][::7] = 0.0
    fr["alpha"][3::7] = 0.001
    fr["alpha"][5::7] = np.nextafter(np.float32(0.001), np.float32(1))
    fr["pixel"][::11] = (fr["pixel"][::11] & 0xffff0000) | np.uint32(s.sc.width + 16 * 7)
    fr["pixel"][1::13] = (fr["pixel"][1::13] & 0xffff) | np.uint32(60000 << 16)
    got, pairs = s.run(fr=fr)
    want, opairs = s.oracle(fr=fr)
    check(got, want)
    assert pairs == opairs
    assert np.all(got[::7] == 0) and np.all(got[3::7] == 0)


def test_forward_lit_entity_out_of_range(ctx, hlib):
    print("[forward-lit] entity indices out of range", flush=True)
    s = Setup(ctx, hlib, "c2_1080p_1k", "clusters")
    ent = s.entity.copy()
    ent[::5] = len(s.transforms) + np.arange(len(ent[::5]), dtype=np.uint32)
    got, pairs = s.run(entity=ent)
    want, opairs = s.oracle(entity=ent)
    check(got, want)
    assert pairs == opairs


def test_forward_lit_no_lights_and_no_fragments(ctx, hlib):
    print("[forward-lit] zero lights / zero fragments", flush=True)
    s = Setup(ctx, hlib, "c1_720p_64", "tiles", num_lights=0)
    got, pairs = s.run(use_transforms=False)
    want, opairs = s.oracle(use_transforms=False)
    assert np.array_equal(got, want) and pairs == opairs == 0  # ambient + emissive, bit for bit
    got0, pairs0 = s.run(fr=s.fr[:0], use_transforms=False)
    assert got0.shape == (0, 4) and pairs0 == 0


def test_forward_lit_4k_16k_sampled(ctx, hlib):
    """The 4K frame's 4.9 M surface fragments against 16384 lights on the cluster grid;
    the oracle checks every 61st fragment (fragments are independent)."""
    print("[forward-lit] 4K/16k clusters, sampled", flush=True)
    s = Setup(ctx, hlib, "c3_4k_16k", "clusters")
    got, pairs = s.run()
    idx = np.arange(0, len(s.fr), 61)
    want, _ = s.oracle(fr=s.fr[idx])
    check(got[idx], want)
    assert pairs > 2 * len(s.fr)
