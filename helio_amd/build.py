"""Builds the native libraries of helio_amd in-tree.

    libhelio_hip.so    HIP kernels + C ABI (hipcc, gfx950)
    libhelio_synth.so  synthetic scene generator (gcc, OpenMP)
    oracle/liboracle.so  CPU checker used by the tests (gcc, -ffp-contract=off)
    tests/cpp/test_passes  C++ tests of include/helio_hip/passes.hpp (g++)

Run as ``python -m helio_amd.build``.  Every call recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
ORACLE = ROOT / "oracle"

HIP_SOURCES = ["hl_api.hip", "hl_cull.hip", "hl_shade.hip", "hl_batch.hip", "hl_gbuffer.hip",
               "hl_forward.hip", "hl_reflect.hip", "hl_interop.hip"]
ARCH = os.environ.get("HELIO_HIP_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=fast-honor-pragmas",
    # hardware FP64 atomic add for the adjoint's per-light accumulators
    "-munsafe-fp-atomics",
    # packed FP32 (v_pk_*_f32) stays off: only scalar VALU FP32 is generated
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{ROOT / 'include'}",
]

# x86-64-v3 (AVX2/FMA/F16C) runs on every host these libraries travel to;
# -ffp-contract=off keeps each C operation a single IEEE rounding.
CC_FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
            "-fPIC", "-shared", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    lines = [l for l in (res.stdout + res.stderr).splitlines()
             if "packed-fp32-ops" not in l and l.strip()]
    if res.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + "\n".join(lines))
    for l in lines:
        if "warning" in l:
            print(l, file=sys.stderr)


def _hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def build_hip() -> Path:
    out = PKG / "libhelio_hip.so"
    objdir = ROOT / "_build"
    objdir.mkdir(exist_ok=True)
    hipcc = _hipcc()
    srcs = [s for s in HIP_SOURCES if (CSRC / s).exists()]

    def compile_one(src: str) -> Path:
        obj = objdir / (src + ".o")
        _run([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_synth() -> Path:
    out = PKG / "libhelio_synth.so"
    _run(["gcc", *CC_FLAGS, str(CSRC / "synth.c"), "-o", str(out), "-lm"])
    return out


def build_oracle() -> Path:
    out = ORACLE / "liboracle.so"
    _run(["gcc", *CC_FLAGS, f"-I{ROOT / 'include'}", str(ORACLE / "oracle.c"),
          str(ORACLE / "oracle64.c"), str(ORACLE / "gbuffer_pack.c"),
          str(ORACLE / "forward_lit.c"), str(ORACLE / "shadow.c"), str(ORACLE / "rc_gi.c"),
          str(ORACLE / "reflection.c"), "-o", str(out), "-lm"])
    return out


def build_cpp_tests() -> Path:
    """The C++ pass-mirror tests, linked against the three libraries in-tree
    (found at run time through an $ORIGIN-relative rpath)."""
    src = ROOT / "tests" / "cpp" / "test_passes.cpp"
    out = src.with_suffix("")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(src),
          "-o", str(out), f"-L{PKG}", f"-L{ORACLE}", "-lhelio_hip", "-lhelio_synth", "-loracle",
          "-Wl,-rpath,$ORIGIN/../../helio_amd:$ORIGIN/../../oracle"])
    return out


def build_all(verbose: bool = True) -> None:
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_hip), ex.submit(build_synth), ex.submit(build_oracle)]
        outs = [f.result() for f in futs]
    outs.append(build_cpp_tests())
    if verbose:
        for o in outs:
            print(f"built {o.relative_to(ROOT)}")


if __name__ == "__main__":
    build_all()
