"""The band of the first per-pixel decision of the deferred shading loop.

The kernels decide "in range and above the 0.002 cut" (deferred_lighting.wgsl:
1128-1131 and 781) with one comparison of the fast squared distance d2 against
a per-light band (lo, hi) built by decision_band() in hl_device.hpp, and
evaluate the shader's own expressions only for d2 inside the band.  This
restates the band computation and checks it against an IEEE FP32 evaluation of
the shader's expressions (dist = length(L - P), atten = 1 / (dist^2 + 1e-4) *
max(0, 1 - (dist / range)^4), radiance = colour * intensity * atten) for
random lights and light-to-pixel vectors clustered around both edges of the
band: every d2 below lo must be in range and above the cut, every d2 above hi
must fail one of the two.  A band placed at the wrong root, or a dropped
margin, fails here.
"""
import math

import numpy as np

f32 = np.float32


def decision_band(range_, ci, cut=True):
    """decision_band() of hl_device.hpp (FP64, lo rounded down, hi up)."""
    R = float(range_) * float(range_)
    lo, hi = R * (1.0 - 1e-5), R * (1.0 + 1e-5)
    if cut:
        tlo, thi = 0.0, 1e-3
        c = float(ci)
        if c > 2e-7:
            b, cc = 0.002 * R, c - 2e-7
            d2c = R * (2.0 * cc / (b + math.sqrt(b * b + 4.0 * c * cc)))
            if d2c < 1e-2:
                tlo, thi = 0.0, 2.0 * d2c + 1e-3
            else:
                tlo, thi = d2c * (1.0 - 1e-4), d2c * (1.0 + 1e-4)
        lo, hi = min(lo, tlo), min(hi, thi)
    lo32, hi32 = f32(lo), f32(hi)
    if float(lo32) > lo:
        lo32 = np.nextafter(lo32, f32(-np.inf))
    if float(hi32) < hi:
        hi32 = np.nextafter(hi32, f32(np.inf))
    return lo32, hi32


def shader_in(t, range_, p1):
    """dist <= range and not all(radiance < 0.002), IEEE FP32 in the shader's order."""
    tx, ty, tz = (t[:, k].astype(f32) for k in range(3))
    s = tx * tx
    s = s + ty * ty
    s = s + tz * tz
    dist = np.sqrt(s)
    atten = f32(1.0) / (dist * dist + f32(0.0001))
    nd = dist / f32(range_)
    atten = atten * np.maximum(f32(0.0), f32(1.0) - nd * nd * nd * nd)
    m = np.maximum(np.maximum(f32(p1[0]) * atten, f32(p1[1]) * atten), f32(p1[2]) * atten)
    return ~(dist > f32(range_)) & ~(m < f32(0.002))


def fast_d2(t):
    tx, ty, tz = (t[:, k].astype(np.float64) for k in range(3))
    zz = (tz.astype(f32) * tz.astype(f32)).astype(np.float64)
    inner = (ty * ty + zz).astype(f32).astype(np.float64)
    return (tx * tx + inner).astype(f32)


def test_band_decides_like_the_shader():
    rng = np.random.default_rng(11)
    checked = 0
    for trial in range(400):
        range_ = f32(rng.uniform(0.05, 40.0))
        inten = rng.choice([rng.uniform(0.5, 5.0), rng.uniform(1e-6, 1e-2), rng.uniform(50, 5000)])
        color = rng.uniform(0.0, 1.0, 3)
        p1 = (color.astype(f32) * f32(inten)).astype(f32)
        ci = f32(p1.max())
        lo, hi = decision_band(range_, ci)
        for edge in (float(lo), float(hi)):
            if edge <= 0.0:
                continue
            # vectors whose squared length straddles the edge by up to 2e-4 relative
            n = 2000
            d = np.sqrt(edge * (1.0 + rng.uniform(-2e-4, 2e-4, n)))
            u = rng.normal(size=(n, 3))
            u /= np.linalg.norm(u, axis=1, keepdims=True)
            t = (u * d[:, None]).astype(f32)
            d2 = fast_d2(t)
            ok = shader_in(t, range_, p1)
            below, above = d2 < lo, d2 > hi
            assert ok[below].all(), (trial, float(range_), float(ci), "below lo but cut")
            assert not ok[above].any(), (trial, float(range_), float(ci), "above hi but kept")
            checked += int(below.sum() + above.sum())
    assert checked > 500_000


def test_band_edges():
    # the cut lies inside the range for any finite intensity, and the range
    # test alone is r^2 (1 -+ 1e-5)
    for r, c in ((1.0, 5.0), (8.0, 0.5), (8.0, 5000.0), (2.0, 1e-5)):
        lo, hi = decision_band(f32(r), f32(c))
        rlo, rhi = decision_band(f32(r), f32(c), cut=False)
        assert lo <= hi and lo <= rlo and hi <= rhi
        assert abs(float(rlo) / (r * r) - (1 - 1e-5)) < 1e-7 and abs(float(rhi) / (r * r) - (1 + 1e-5)) < 1e-7
    # a light too dim to reach 0.002 anywhere: only d2 <= 1e-3 is left to the exact test
    lo, hi = decision_band(f32(5.0), f32(1e-7))
    assert lo == 0.0 and abs(float(hi) - 1e-3) < 1e-9
