"""GPU adjoint vs the FP64 oracle and central finite differences.

Workload: 512x512 fp32 G-buffer (and 1024x1024 for the FP64 comparison), 256
point lights, target rendered from the
ground truth, guess perturbed (albedo N(0,0.1), roughness N(0,0.1) clamped to
[0.05,1], light position N(0,0.5), intensity x U[0.7,1.3]); per-pixel L2 loss;
clustered lists (16x9x24, uncapped, so every pixel sees all its lights).

Bars:
  * gradient element vs FP64 analytic: |hip - f64| <= 1e-3 |f64| + 1e-6 * scale,
    scale = largest magnitude of that gradient kind (a floor for elements that
    cancel to ~0);
  * 1000 sampled parameters vs central FD of the FP64 forward (h = 1e-3):
    same bar, stencils crossing a cut-off skipped;
  * forward radiance vs FP32 restatement: 1e-4 rel + 1e-6; loss vs FP64: 1e-4 rel.
"""
import ctypes as C

import numpy as np
import pytest

from helio_amd import abi, synth
from helio_amd.device import ClusterGrid, LightCullPass, upload_scene
import adjoint_util as A
import oracle_lib as O

pytestmark = pytest.mark.gpu


def make_prob(size):
    sc, target, _ = A.make_problem(size, size, 256)
    px = A.decode(sc)
    dpx64, dpos64, dint64 = A.backward64(sc, px, sc.lights, target)
    _, loss64 = A.forward64(sc, px, sc.lights, target)
    return sc, target, px, dpx64, dpos64, dint64, loss64


@pytest.fixture(scope="module")
def prob():
    return make_prob(512)


# 1024x1024: 4096 block tiles of 32x8 pixels, where the fused launch gives
# each wave a tile of its own (512x512 shares each wave tile among four waves)
@pytest.fixture(scope="module")
def prob1k():
    return make_prob(1024)


def run_hip(ctx, hlib, sc, target, fused=True):
    W, H = sc.width, sc.height
    ds = upload_scene(ctx, sc)
    params = synth.cluster_params(hlib, sc, cap=0)
    cull = LightCullPass(ctx, W, H, clusters=params, pool_cap=1 << 22)
    cull.prepare(ds)
    cull.execute(ds)
    grid = cull.publish()
    tgt = ctx.upload(target)
    out = ctx.alloc(W * H * 16)
    loss = ctx.alloc(8)
    dpx = ctx.alloc(5 * W * H * 4)
    n = len(sc.lights)
    dpos = ctx.alloc(3 * n * 4)
    dint = ctx.alloc(n * 4)
    g = sc.globals
    L = ctx.lib
    L.call("hl_deferred_forward_backward_l2", ctx.h, ds.camera.ptr, C.byref(g), C.byref(ds.gbuffer),
           ds.lights.ptr, n, C.byref(grid), tgt.ptr, W * 16, out.ptr, W * 16, loss.ptr, dpx.ptr, dpos.ptr,
           dint.ptr)
    res = dict(out=out.download(np.float32, (H, W, 4)), loss=float(loss.download(np.float64)[0]),
               dpx=dpx.download(np.float32, (5, H, W)), dpos=dpos.download(np.float32, (n, 3)),
               dint=dint.download(np.float32))
    if not fused:
        return res
    # plain backward with dL/dout = 2 (out - target)
    dout = np.zeros((H, W, 4), np.float32)
    dout[..., :3] = 2.0 * (res["out"][..., :3] - target[..., :3])
    dout[sc.depth >= 1.0] = 0.0
    dd = ctx.upload(dout)
    L.call("hl_deferred_backward", ctx.h, ds.camera.ptr, C.byref(g), C.byref(ds.gbuffer), ds.lights.ptr, n,
           C.byref(grid), dd.ptr, W * 16, dpx.ptr, dpos.ptr, dint.ptr)
    res["b_dpx"] = dpx.download(np.float32, (5, H, W))
    res["b_dpos"] = dpos.download(np.float32, (n, 3))
    res["b_dint"] = dint.download(np.float32)
    return res


def within(hip, ref, scale):
    err = np.abs(hip.astype(np.float64) - ref)
    tol = 1e-3 * np.abs(ref) + 1e-6 * scale
    return err <= tol, err / (np.abs(ref) + 1e-6 * scale)


@pytest.mark.parametrize("which", ["prob", "prob1k"])
def test_adjoint_matches_fp64(ctx, hlib, request, which):
    sc, target, px, dpx64, dpos64, dint64, loss64 = request.getfixturevalue(which)
    print(f"[adjoint] {sc.width}x{sc.height}, 256 lights", flush=True)
    r = run_hip(ctx, hlib, sc, target)
    sky = sc.depth >= 1.0
    # forward part against the FP32 restatement over the same (all-lights) semantics
    n = len(sc.lights)
    p = abi.hl_cluster_params()
    p.grid_x = p.grid_y = p.grid_z = 1
    p.cell_w, p.cell_h, p.width, p.height = sc.width, sc.height, sc.width, sc.height
    off0, cnt0, pool0 = np.zeros(1, np.uint32), np.full(1, n, np.uint32), np.arange(n, dtype=np.uint32)
    ref, _ = O.shade(sc, O.cluster_grid_host(p, off0, cnt0, pool0))
    e = np.abs(r["out"][~sky] - ref[~sky])
    assert np.all(e <= 1e-4 * np.abs(ref[~sky]) + 1e-6)
    assert abs(r["loss"] - loss64) <= 1e-4 * loss64
    for k in range(5):
        ok, rel = within(r["dpx"][k], dpx64[k], np.abs(dpx64[k]).max())
        print(f"  pixel grad {k}: max scaled rel {rel.max():.2e}, frac bad {1 - ok.mean():.2e}")
        assert ok.all(), k
    ok, rel = within(r["dpos"], dpos64, np.abs(dpos64).max())
    print(f"  light pos grad: max scaled rel {rel.max():.2e}")
    assert ok.all()
    ok, rel = within(r["dint"], dint64, np.abs(dint64).max())
    print(f"  light intensity grad: max scaled rel {rel.max():.2e}")
    assert ok.all()
    # the plain backward with dL/dout = 2(out - target) gives the same gradients
    for a, b in (("dpx", "b_dpx"), ("dpos", "b_dpos"), ("dint", "b_dint")):
        sc_ = np.abs(r[a]).max()
        assert np.all(np.abs(r[a] - r[b]) <= 1e-5 * np.abs(r[a]) + 1e-6 * sc_), a


def test_adjoint_matches_finite_differences(ctx, hlib, prob):
    sc, target, px, dpx64, dpos64, dint64, _ = prob
    print("[adjoint FD] 1000 sampled parameters", flush=True)
    r = run_hip(ctx, hlib, sc, target, fused=False)
    rng = np.random.default_rng(2024)
    lit = np.nonzero(px["sky"] == 0)[0]
    samples = [(int(k), int(i)) for k, i in zip(rng.integers(0, 5, 800), rng.choice(lit, 800, replace=False))]
    samples += [(int(k), int(i)) for k, i in zip(rng.integers(5, 9, 200), rng.integers(0, len(sc.lights), 200))]
    scale = {k: np.abs(dpx64[k]).max() for k in range(5)}
    scale.update({k: np.abs(dpos64).max() for k in (5, 6, 7)})
    scale[8] = np.abs(dint64).max()
    checked = bad = 0
    worst = 0.0
    for kind, idx in samples:
        fd, cross = A.fd64(sc, px, sc.lights, target, kind, idx)
        if cross:
            continue
        hip = (r["dpx"][kind].reshape(-1)[idx] if kind < 5 else
               (r["dpos"][idx, kind - 5] if kind < 8 else r["dint"][idx]))
        err = abs(float(hip) - fd)
        tol = 1e-3 * abs(fd) + 1e-6 * scale[kind]
        worst = max(worst, err / (abs(fd) + 1e-6 * scale[kind]))
        checked += 1
        bad += err > tol
    print(f"  checked {checked} of 1000 (others straddle a cut-off), worst scaled rel {worst:.2e}")
    assert checked >= 900
    assert bad == 0
