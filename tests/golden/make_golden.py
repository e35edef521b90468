"""Writes the golden fixtures of the deferred-lighting path.

Each fixture holds small inputs (camera, lights, G-buffer planes, cluster
parameters) and the outputs of the CPU restatement on them: tile lists and
counts, cluster offsets/counts/pool, radiance over tiles and clusters, and Q.
The reference (Rust + WGSL on wgpu) holds no vectors for this path and cannot
run here, so these are the oracle's outputs, frozen: tests check that the
oracle still reproduces them bit for bit and that the HIP path matches them.

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from helio_amd import abi, synth  # noqa: E402
import oracle_lib as O  # noqa: E402

FIXTURES = {
    # name: (workload, overrides, cluster grid, cap)
    "tiny_ref_formats": ("c2_1080p_1k", dict(width=96, height=64, num_lights=96), (6, 4, 8), 0),
    "tiny_ref_capped": ("c2_1080p_1k", dict(width=80, height=48, num_lights=160), (5, 3, 6), 12),
    "tiny_fp16": ("c3_4k_16k", dict(width=128, height=72, num_lights=512), (8, 4, 12), 0),
    "tiny_fp32": ("c4_512_adj", dict(width=64, height=64, num_lights=48), (4, 4, 8), 0),
}


def cluster_params_py(scene, grid, cap):
    """Fixture cluster params come from the library's host helper (no GPU)."""
    lib = abi.Lib()
    return synth.cluster_params(lib, scene, grid=grid, cap=cap)


def with_directional(scene):
    lights = scene.lights.copy()
    if len(lights) > 3:
        lights[3]["light_type"] = abi.HL_LIGHT_DIRECTIONAL
        lights[3]["direction_outer"][:3] = (0.3, -0.9, 0.3)
        lights[3]["color_intensity"][3] = 0.25
    scene.lights = lights
    return scene


def compute(scene, params):
    tl, tc = O.cull_tiles(scene)
    flat = np.ascontiguousarray(tl.reshape(-1))
    out_t, q_t = O.shade(scene, O.tile_grid_host(scene, flat, tc))
    cap = 1 << 20
    off, cnt, pool, flags = O.cull_clusters(scene, params, cap)
    used = int(off[-1] + cnt[-1])
    pool = pool[:used].copy()
    out_c, q_c = O.shade(scene, O.cluster_grid_host(params, off, cnt, pool))
    return dict(tile_lists=tl, tile_counts=tc, radiance_tiles=out_t, q_tiles=np.uint64(q_t),
                cl_offsets=off, cl_counts=cnt, cl_pool=pool, cl_flags=np.uint32(flags),
                radiance_clusters=out_c, q_clusters=np.uint64(q_c))


def inputs(scene, params):
    return dict(camera=np.frombuffer(bytes(scene.camera), np.uint8).copy(),
                globals=np.frombuffer(bytes(scene.globals), np.uint8).copy(),
                cluster_params=np.frombuffer(bytes(params), np.uint8).copy(),
                lights=scene.lights.view(np.uint8).copy(), albedo=scene.albedo, normal=scene.normal,
                orm=scene.orm, emissive=scene.emissive, depth=scene.depth,
                formats=np.array(scene.formats, np.uint32), size=np.array([scene.width, scene.height], np.uint32))


def main():
    for name, (work, over, grid, cap) in FIXTURES.items():
        sc = with_directional(synth.make_scene(work, **over))
        params = cluster_params_py(sc, grid, cap)
        data = inputs(sc, params)
        data.update(compute(sc, params))
        np.savez_compressed(HERE / f"{name}.npz", **data)
        print(f"wrote {name}.npz: Q tiles {int(data['q_tiles'])}, clusters {int(data['q_clusters'])}")


if __name__ == "__main__":
    main()
