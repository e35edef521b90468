"""include/helio_hip/roofline.h is the one place the algorithmic-work
constants and peaks live: it compiles as C, and bench.py reads the same values."""
import importlib.util
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
NAMES = ["HL_RL_SKY_PIXEL_BYTES", "HL_RL_DEPTH_BYTES", "HL_RL_OUTPUT_BYTES", "HL_RL_CELL_BYTES",
         "HL_RL_LIST_ENTRY_BYTES", "HL_RL_LIGHT_BYTES", "HL_RL_FLOP_PER_PIXEL", "HL_RL_FLOP_PER_PAIR",
         "HL_PEAK_HBM_GBS", "HL_PEAK_FP32_TFLOPS", "HL_PEAK_FP32_SCALAR_TFLOPS"]


def test_header_compiles_and_matches_bench(tmp_path):
    src = tmp_path / "rl.c"
    src.write_text('#include <stdio.h>\n#include "helio_hip/roofline.h"\nint main(void) {\n'
                   + "".join(f'    printf("%s %.6f\\n", "{n}", (double)({n}));\n' for n in NAMES)
                   + "    return 0;\n}\n")
    exe = tmp_path / "rl"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    compiled = {k: float(v) for k, v in (line.split() for line in out.splitlines())}
    spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for n in NAMES:
        assert abs(bench.RL[n] - compiled[n]) < 1e-9, n
    # the 4K figures of the roofline: 52 B per shaded pixel at fp16 (4 x 8 + 4 + 16)
    assert 4 * 8 + compiled["HL_RL_DEPTH_BYTES"] + compiled["HL_RL_OUTPUT_BYTES"] == 52
