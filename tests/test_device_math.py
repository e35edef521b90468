"""Exactness claims of the shading kernels' integer-free arithmetic, checked on
the CPU in exact arithmetic (no GPU needed):

  * hl_unorm8: fma(v, hi, v * lo) with hi = float(1/255), lo = float(1/255 - hi)
    equals the IEEE quotient v / 255 for every v in 0..255 -- the constants in
    hl_device.hpp are the ones derived here;
  * hl_floor_div: floor(fma(x, r, 0.5 r)) with r = 1/c rounded to within one
    ulp either way equals x // c for every x the kernels pass (< 2^15 + 256)
    and every cell size up to 1024 plus a spread of larger ones.
"""
import re
from fractions import Fraction
from pathlib import Path

import numpy as np

HDR = Path(__file__).resolve().parents[1] / "helio_amd" / "csrc" / "hl_device.hpp"


def rn32(x: Fraction) -> np.float32:
    """Round a rational to the nearest float32, ties to even."""
    f = np.float32(float(x))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    return min(cands, key=lambda c: (abs(Fraction(float(c)) - x), int(c.view(np.uint32)) & 1))


def test_unorm8_constants_and_exactness():
    src = HDR.read_text()
    hi_s = re.search(r"const float hi = ([0-9.e+-]+)f;", src).group(1)
    lo_s = re.search(r"const float lo = ([0-9.e+-]+)f;", src).group(1)
    hi = rn32(Fraction(1, 255))
    lo = rn32(Fraction(1, 255) - Fraction(float(hi)))
    assert np.float32(hi_s) == hi and np.float32(lo_s) == lo
    for v in range(256):
        prod_lo = rn32(Fraction(v) * Fraction(float(lo)))
        q = rn32(Fraction(v) * Fraction(float(hi)) + Fraction(float(prod_lo)))
        assert q == rn32(Fraction(v, 255)), v


def test_floor_div_float_estimate():
    x = np.arange(0, (1 << 15) + 256, dtype=np.float64)
    cells = list(range(1, 1025)) + [1031, 1500, 2048, 4097, 8191, 16384, 32768]
    for c in cells:
        r0 = np.float32(1.0) / np.float32(c)
        want = np.floor_divide(x, c)
        for r in (np.nextafter(r0, np.float32(0)), r0, np.nextafter(r0, np.float32(2))):
            # x * r + 0.5 r is exact in float64 (<= 40 significant bits), so one
            # conversion to float32 is the single rounding of the fma
            est = (x * np.float64(r) + 0.5 * np.float64(r)).astype(np.float32)
            assert np.array_equal(np.floor(est.astype(np.float64)), want), c
