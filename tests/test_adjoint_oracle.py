"""Pins the FP64 analytic adjoint of the oracle against central finite
differences of the FP64 forward (h = 1e-3), on sampled per-pixel material and
per-light parameters, skipping samples whose stencil crosses a cut-off
(range, NdL = 0, the 0.002 radiance cut-off, F0 clamp)."""
import numpy as np
import pytest

import adjoint_util as A


@pytest.fixture(scope="module")
def problem():
    sc, target, _ = A.make_problem(width=96, height=96, num_lights=64)
    px = A.decode(sc)
    return sc, px, target


def rel_err(a, b, scale):
    return abs(a - b) / max(abs(a), abs(b), scale)


def test_loss_is_positive_and_gradient_nonzero(problem):
    sc, px, target = problem
    _, loss = A.forward64(sc, px, sc.lights, target)
    assert loss > 0
    dpx, dpos, dint = A.backward64(sc, px, sc.lights, target)
    assert np.abs(dpx).max() > 0 and np.abs(dpos).max() > 0 and np.abs(dint).max() > 0


def test_pixel_gradients_match_fd(problem):
    sc, px, target = problem
    dpx, _, _ = A.backward64(sc, px, sc.lights, target)
    rng = np.random.default_rng(1)
    lit = np.nonzero(px["sky"] == 0)[0]
    scale = 1e-3 * np.abs(dpx).max()
    checked = 0
    for i in rng.choice(lit, 120, replace=False):
        for kind in range(5):
            fd, cross = A.fd64(sc, px, sc.lights, target, kind, int(i))
            if cross:
                continue
            an = dpx[kind].reshape(-1)[i]
            assert rel_err(an, fd, scale) < 1e-3, (kind, int(i), an, fd)
            checked += 1
    assert checked > 400


def test_light_gradients_match_fd(problem):
    sc, px, target = problem
    _, dpos, dint = A.backward64(sc, px, sc.lights, target)
    scale_p = 1e-3 * np.abs(dpos).max()
    scale_i = 1e-3 * np.abs(dint).max()
    checked = 0
    for li in range(len(sc.lights)):
        for kind in range(5, 9):
            fd, cross = A.fd64(sc, px, sc.lights, target, kind, li)
            if cross:
                continue
            an = dpos[li, kind - 5] if kind < 8 else dint[li]
            assert rel_err(an, fd, scale_p if kind < 8 else scale_i) < 1e-3, (kind, li, an, fd)
            checked += 1
    assert checked > 30


def test_ambient_only_gradient_closed_form():
    """No lights: out = mix(0.15 s, s, 0.5 N.y + 0.5) a ao, so dL/da = 2 (out - t) mix ao."""
    sc, target, _ = A.make_problem(width=32, height=32, num_lights=1)
    sc.lights = sc.lights[:0].copy()
    px = A.decode(sc)
    out, _ = A.forward64(sc, px, sc.lights, target)
    dpx, _, _ = A.backward64(sc, px, sc.lights, target)
    s = np.array(sc.globals.ambient_color[:3], dtype=np.float64) * float(sc.globals.ambient_intensity)
    ny = px["N"][:, 1]
    t = 0.5 * ny + 0.5
    mixv = 0.15 * s[None, :] * (1 - t[:, None]) + s[None, :] * t[:, None]
    want = 2 * (out.reshape(-1, 3) - target[..., :3].reshape(-1, 3)) * mixv * px["ao"][:, None]
    want[px["sky"] == 1] = 0
    np.testing.assert_allclose(dpx[:3].reshape(3, -1).T, want, rtol=1e-12, atol=1e-15)
    assert np.all(dpx[3:] == 0)
