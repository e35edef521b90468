"""CPU tests of the boundary: record layouts, exported symbols, host helpers."""
import ctypes as C
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from helio_amd import abi

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "helio_hip" / "helio_hip.h"


def test_record_sizes_match_reference_asserts():
    # crates/libhelio/src/light.rs:111 (128), camera.rs layout (368),
    # helio-pass-light-cull/src/lib.rs:24-35 (32), deferred-light/src/lib.rs:15-52 (112),
    # material.rs:231-254 (96)
    assert abi.GPU_LIGHT_DTYPE.itemsize == 128
    assert C.sizeof(abi.GpuCameraUniforms) == 368
    assert C.sizeof(abi.LightCullParams) == 32
    assert C.sizeof(abi.DeferredGlobals) == 112
    assert C.sizeof(abi.GpuMaterial) == 96


@pytest.mark.parametrize("field,off", [("position_range", 0), ("direction_outer", 16),
                                       ("color_intensity", 32), ("shadow_index", 48),
                                       ("light_type", 52), ("inner_angle", 56),
                                       ("god_rays_enabled", 64), ("flare_enabled", 84),
                                       ("flare_scale", 96), ("ies_profile_index", 112),
                                       ("light_function_index", 116), ("ies_angle_offset", 124)])
def test_gpulight_offsets(field, off):
    assert abi.GPU_LIGHT_DTYPE.fields[field][1] == off


@pytest.mark.parametrize("field,off", [("view", 0), ("proj", 64), ("view_proj", 128),
                                       ("inv_view_proj", 192), ("position_near", 256),
                                       ("forward_far", 272), ("jitter_frame", 288),
                                       ("prev_view_proj", 304)])
def test_camera_offsets(field, off):
    assert getattr(abi.GpuCameraUniforms, field).offset == off


@pytest.mark.parametrize("field,off", [("ambient_color", 16), ("csm_splits", 64), ("debug_mode", 80),
                                       ("has_rc_gi", 84), ("num_tiles_x", 88),
                                       ("enable_env_reflections", 100)])
def test_globals_offsets(field, off):
    assert getattr(abi.DeferredGlobals, field).offset == off


def test_c_header_static_asserts_compile(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "helio_hip/helio_hip.h"\nint main(void){return 0;}\n')
    r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", f"-I{ROOT / 'include'}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    srcpp = tmp_path / "t.cpp"
    srcpp.write_text('#include "helio_hip/helio_hip.h"\nint main(){return 0;}\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{ROOT / 'include'}", str(srcpp)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_ctypes_structs_match_c_sizes(tmp_path):
    src = tmp_path / "s.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "helio_hip/helio_hip.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n",'
                   'sizeof(hl_gbuffer), sizeof(hl_cluster_params), sizeof(hl_light_grid), sizeof(hl_view_desc),'
                   'sizeof(hl_texture), sizeof(hl_surface), sizeof(hl_fragment), sizeof(hl_transform),'
                   'sizeof(ForwardLitGlobals), sizeof(hl_shadow_maps), sizeof(GpuReflectionCapture),'
                   'sizeof(hl_cube_array), sizeof(hl_reflection_inputs),'
                   'offsetof(hl_reflection_inputs, rc_cascade0));return 0;}\n')
    exe = tmp_path / "s"
    subprocess.run(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert sizes == [C.sizeof(abi.hl_gbuffer), C.sizeof(abi.hl_cluster_params),
                     C.sizeof(abi.hl_light_grid), C.sizeof(abi.hl_view_desc), C.sizeof(abi.hl_texture),
                     C.sizeof(abi.hl_surface), C.sizeof(abi.hl_fragment), C.sizeof(abi.hl_transform),
                     C.sizeof(abi.ForwardLitGlobals), C.sizeof(abi.hl_shadow_maps),
                     C.sizeof(abi.GpuReflectionCapture), C.sizeof(abi.hl_cube_array),
                     C.sizeof(abi.hl_reflection_inputs), abi.hl_reflection_inputs.rc_cascade0.offset]


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"HL_API\s+[^;(]*?\b(hl_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) > 20
    lib = C.CDLL(str(abi.library_path()))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers the header
    assert set(names) <= set(abi.SIGNATURES), set(names) - set(abi.SIGNATURES)


def test_version_and_status_strings(hlib):
    assert b"gfx950" in hlib.raw["hl_version"]()
    assert hlib.raw["hl_status_string"](abi.HL_EINVAL) == b"invalid argument"


def test_context_create_fails_cleanly_without_gpu(hlib):
    import os
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    st = hlib.raw["hl_context_create"](0, C.byref(h))
    assert st in (abi.HL_EHIP, abi.HL_EINVAL) and not h.value


def test_null_context_is_einval(hlib):
    assert hlib.raw["hl_synchronize"](None) == abi.HL_EINVAL
    assert hlib.raw["hl_light_cull_tiles"](None, None, None, None, None, None) == abi.HL_EINVAL
    assert hlib.raw["hl_deferred_forward"](None, None, None, None, None, 0, None, None, 0, 0) == abi.HL_EINVAL


def test_cluster_params_closed_form(hlib):
    from helio_amd import synth
    s = synth.spec("c3_4k_16k")
    cam = synth.make_camera(s)
    p = abi.hl_cluster_params()
    hlib.call("hl_cluster_params_init", C.byref(cam), 3840, 2160, 32, 18, 48, 0, C.byref(p))
    assert (p.cell_w, p.cell_h) == (120, 120)
    near, far = 0.1, 1000.0
    p22, p32 = float(cam.proj[10]), float(cam.proj[14])
    split = np.array(p.depth_split[:49], dtype=np.float64)
    assert split[0] == 0.0 and split[48] == 1.0
    assert np.all(np.diff(split) > 0)
    for k in range(1, 48):
        z = near * (far / near) ** (k / 48)      # log slicing: z_k = near (far/near)^(k/S)
        d = (p22 * (-z) + p32) / z               # NDC depth of view distance z (standard Z)
        assert abs(split[k] - d) <= 2 * np.spacing(np.float32(d)), k
    # the padded view-distance ranges contain the unpadded ones and overlap
    for k in range(48):
        z0 = near * (far / near) ** (k / 48)
        z1 = near * (far / near) ** ((k + 1) / 48)
        assert p.slice_near[k] < z0 * (1 + 1e-6) and p.slice_far[k] > z1 * (1 - 1e-6)
        assert p.slice_r_pad[k] > 0
    assert p.r_pad_column == pytest.approx(max(p.slice_r_pad[:48]))


def test_cluster_params_rejects_non_perspective(hlib):
    cam = abi.GpuCameraUniforms()
    for i in (0, 5, 10, 15):
        cam.proj[i] = 1.0  # orthographic-like
    cam.position_near[3] = 0.1
    cam.forward_far[3] = 100.0
    p = abi.hl_cluster_params()
    st = hlib.raw["hl_cluster_params_init"](C.byref(cam), 64, 64, 4, 4, 4, 0, C.byref(p))
    assert st == abi.HL_EUNSUPPORTED


def test_header_constants_match_python_mirror():
    """Every numeric #define of helio_hip.h that abi.py mirrors has the same value."""
    seen = 0
    text = HEADER.read_text()
    for name, val in re.findall(r"^#define\s+(HL_[A-Z0-9_]+)\s+\(?(-?(?:0x[0-9a-fA-F]+|\d+))u?\)?\s*(?:/\*.*)?$",
                                text, re.M):
        if hasattr(abi, name):
            assert getattr(abi, name) == int(val, 0), name
            seen += 1
    assert seen >= 12
    m = re.search(r"#define HL_NOALLOC_SCRATCH_BYTES \((\d+)u \+ (\d+)u \* HL_NOALLOC_MAX_LIGHTS\)", text)
    assert m and int(m.group(1)) + int(m.group(2)) * abi.HL_NOALLOC_MAX_LIGHTS == abi.HL_NOALLOC_SCRATCH_BYTES


def _c_exports():
    """name -> parameter count of every HL_API function in helio_hip.h."""
    out = {}
    for m in re.finditer(r"HL_API\s+[\w\s\*]+?\b(hl_\w+)\s*\(([^;]*?)\)\s*;", HEADER.read_text(), re.S):
        args = " ".join(m.group(2).split())
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_rust_binding_declares_every_export():
    """The extern block a maintainer adds on the reference side
    declares every exported function with the same number of parameters."""
    text = (ROOT / "INTEGRATION.md").read_text()
    block = text[text.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    rust = {}
    for m in re.finditer(r"pub fn (hl_\w+)\(([^)]*)\)", block, re.S):
        args = " ".join(m.group(2).split())
        rust[m.group(1)] = 0 if not args else args.count(",") + 1
    c = _c_exports()
    assert len(c) >= 38
    assert sorted(set(c) - set(rust)) == [], "declared in C, missing in the Rust binding"
    assert sorted(set(rust) - set(c)) == [], "in the Rust binding, not exported"
    for name, n in c.items():
        assert rust[name] == n, (name, n, rust[name])
