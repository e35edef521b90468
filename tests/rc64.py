"""sample_rc_irradiance (deferred_lighting.wgsl:824-904) restated in float64,
vectorised over points, written from the shader's description independently
of oracle/rc_gi.c: the volume fade (a 5 % smoothstep margin on each axis of
the probe box), cosine weights of the normal against the 16 octahedral probe
directions (rc_oct_decode at the 4 x 4 bin centres), the probe cell of the
point (probe centres at (i + 0.5) cells), per corner probe the weighted mean
of its direction texels with weight > 0.001 (divided by max(sum, 0.001)),
and the trilinear blend of the 8 corner probes (z, then y, then x).
Test infrastructure only."""
import numpy as np

PROBES, DIRS = 16, 4


def oct_dirs64():
    out = np.zeros((DIRS * DIRS, 3))
    for dx in range(DIRS):
        for dy in range(DIRS):
            fx, fy = (dx + 0.5) / DIRS * 2 - 1, (dy + 0.5) / DIRS * 2 - 1
            l1 = abs(fx) + abs(fy)
            if l1 > 1:
                n = np.array([(1 - abs(fy)) * (1 if fx >= 0 else -1), 1 - l1, (1 - abs(fx)) * (1 if fy >= 0 else -1)])
            else:
                n = np.array([fx, 1 - l1, fy])
            out[dx * DIRS + dy] = n / np.linalg.norm(n)
    return out


def smoothstep(e0, e1, x):
    q = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return q * q * (3 - 2 * q)


def rc_irradiance64(atlas, wmin, wmax, wp, N, texel_order="xy", fade=True):
    """atlas: [1024, 64, 4] (float16/32); wp, N: [n, 3].  Returns (irradiance
    [n, 3], weights [n, 16]).  texel_order / fade: planted mistakes for tests
    ("yx" swaps the direction bin's texel axes; fade=False drops the fade)."""
    tex = np.asarray(atlas, np.float64)[..., :3]
    wmin, wmax = np.asarray(wmin, np.float64), np.asarray(wmax, np.float64)
    wp, N = np.asarray(wp, np.float64), np.asarray(N, np.float64)
    size = wmax - wmin
    n = len(wp)
    if np.any(size <= 0):
        return np.zeros((n, 3)), np.zeros((n, DIRS * DIRS))
    t = (wp - wmin) / size
    vw = np.prod(smoothstep(0.0, 0.05, t) * smoothstep(1.0, 0.95, t), axis=1) if fade else np.ones(n)
    cw = np.maximum(0.0, N @ oct_dirs64().T)  # [n, 16], bin index dx * 4 + dy
    pf = np.clip((wp - wmin) / (size / PROBES) - 0.5, 0.0, PROBES - 1.0)
    pi = np.floor(pf).astype(np.int64)
    fr = pf - pi
    use = cw > 0.001
    wsum = np.maximum((cw * use).sum(1), 0.001)
    dx, dy = np.divmod(np.arange(DIRS * DIRS), DIRS)
    if texel_order == "yx":
        dx, dy = dy, dx
    corner = np.zeros((8, n, 3))
    for b in range(8):
        p = np.minimum(pi + np.array([b >> 2, (b >> 1) & 1, b & 1]), PROBES - 1)
        rows = ((p[:, 1] * PROBES + p[:, 2]) * DIRS)[:, None] + dy[None, :]
        cols = (p[:, 0] * DIRS)[:, None] + dx[None, :]
        c = tex[rows, cols]  # [n, 16, 3]
        corner[b] = ((cw * use)[..., None] * c).sum(1) / wsum[:, None]

    def mix(a, b, f):
        return a * (1 - f[:, None]) + b * f[:, None]

    c0 = mix(mix(corner[0], corner[1], fr[:, 2]), mix(corner[2], corner[3], fr[:, 2]), fr[:, 1])
    c1 = mix(mix(corner[4], corner[5], fr[:, 2]), mix(corner[6], corner[7], fr[:, 2]), fr[:, 1])
    out = mix(c0, c1, fr[:, 0]) * vw[:, None]
    out[vw <= 0] = 0.0
    return out, cw


def world_pos(sc, xs, ys):
    """The shader's world positions of pixels (FP32, oracle.c or_world_pos)."""
    import ctypes as C
    import oracle_lib as O
    out = np.zeros((len(xs), 3))
    o = (C.c_float * 3)()
    for i, (x, y) in enumerate(zip(xs, ys)):
        O.oracle().or_world_pos(C.byref(sc.camera), sc.width, sc.height, int(x), int(y), C.c_float(sc.depth[y, x]), o)
        out[i] = o[:]
    return out


def ambient64(sc, atlas, lo, hi, xs, ys):
    """The ambient + emissive of pixels (xs, ys) of an fp32-G-buffer scene with
    radiance-cascade GI (deferred_lighting.wgsl:1159-1163, 1189-1231) in
    float64: mix(hemi, kD_ibl rc albedo, clamp(4 |rc|, 0, 1)) ao + emissive,
    kD_ibl = (1 - F) (1 - metallic), F = F0 + (max(1 - r, F0) - F0) (1 - NdV)^5,
    hemi = mix(sky 0.15, sky, N.y 0.5 + 0.5) albedo.  Returns (rgb [n, 3],
    cosine weights [n, 16])."""
    H, W = sc.height, sc.width
    view = lambda a: a.view(np.float32).reshape(H, W, 4).astype(np.float64)[ys, xs]  # noqa: E731
    alb, nrm, orm, emi = view(sc.albedo), view(sc.normal), view(sc.orm), view(sc.emissive)
    wp = world_pos(sc, xs, ys)
    N = nrm[:, :3] / np.linalg.norm(nrm[:, :3], axis=1, keepdims=True)
    rc, cw = rc_irradiance64(atlas, lo, hi, wp, N)
    eye = np.array(sc.camera.position_near[:3], np.float64)
    V = (eye - wp) / np.linalg.norm(eye - wp, axis=1, keepdims=True)
    ndv = np.maximum((N * V).sum(1), 0.0)
    F0 = np.clip(np.stack([nrm[:, 3], orm[:, 3], emi[:, 3]], 1), 0.0, 0.999)
    r, m, ao = orm[:, 1:2], orm[:, 2:3], orm[:, 0:1]
    F = F0 + (np.maximum(1 - r, F0) - F0) * (1 - ndv[:, None]) ** 5
    diff = (1 - F) * (1 - m) * rc * alb[:, :3]
    sky = np.array(sc.globals.ambient_color[:3], np.float64) * sc.globals.ambient_intensity
    t = N[:, 1:2] * 0.5 + 0.5
    hemi = (sky * 0.15 * (1 - t) + sky * t) * alb[:, :3]
    w = np.clip(np.linalg.norm(rc, axis=1, keepdims=True) * 4, 0.0, 1.0)
    return (hemi * (1 - w) + diff * w) * ao + emi[:, :3], cw
