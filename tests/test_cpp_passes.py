"""The C++ pass mirror (include/helio_hip/passes.hpp), exercised by the compiled
driver tests/cpp/test_passes (built with the libraries by helio_amd/build.py):

  * on the CPU: error behaviour without a GPU, cluster parameters, constants;
  * on the GPU: LightCull (tiles, clusters) and DeferredLight driven through the
    pass objects as a render graph would, lists bit-exact and radiance within
    1e-4 relative of the oracle, errors surfacing as exceptions.
"""
import subprocess
from pathlib import Path

import pytest

DRIVER = Path(__file__).resolve().parent / "cpp" / "test_passes"


def run(mode: str) -> None:
    if not DRIVER.exists():
        pytest.fail(f"{DRIVER} is not built: run `python -m helio_amd.build`")
    res = subprocess.run([str(DRIVER), mode], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "FAIL" not in res.stdout


def test_cpp_passes_cpu():
    run("cpu")


@pytest.mark.gpu
def test_cpp_passes_gpu():
    run("gpu")
