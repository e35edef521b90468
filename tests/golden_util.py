"""Loads the golden fixtures under tests/golden/ into HostScene objects."""
import ctypes as C
from pathlib import Path

import numpy as np

from helio_amd import abi, synth

GOLDEN = Path(__file__).resolve().parent / "golden"


def names():
    return sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    cam = abi.GpuCameraUniforms.from_buffer_copy(z["camera"].tobytes())
    glb = abi.DeferredGlobals.from_buffer_copy(z["globals"].tobytes())
    params = abi.hl_cluster_params.from_buffer_copy(z["cluster_params"].tobytes())
    lights = z["lights"].view(abi.GPU_LIGHT_DTYPE).copy()
    W, H = (int(v) for v in z["size"])
    fa, ta, fn, tn = (int(v) for v in z["formats"])
    spec = synth.SceneSpec(name, W, H, 0, len(lights))
    sc = synth.HostScene(spec=spec, camera=cam, lights=lights, globals=glb,
                         albedo=np.ascontiguousarray(z["albedo"]), normal=np.ascontiguousarray(z["normal"]),
                         orm=np.ascontiguousarray(z["orm"]), emissive=np.ascontiguousarray(z["emissive"]),
                         depth=np.ascontiguousarray(z["depth"]), formats=(fa, ta, fn, tn), texel=(ta, tn))
    expected = {k: z[k] for k in z.files if k not in ("camera", "globals", "cluster_params", "lights", "albedo",
                                                     "normal", "orm", "emissive", "depth", "formats", "size")}
    return sc, params, expected
