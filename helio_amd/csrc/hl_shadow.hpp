// hl_shadow.hpp -- shadow_factor of the deferred shader
// (crates/passes/3d/helio-pass-deferred-light/shaders/deferred_lighting.wgsl:
// 234-565) for the shading kernel: normal-offset bias, the light's layer
// (cube face, spot, or directional cascade with its blend zone), then PCF
// over a Vogel disk or PCSS, each tap the minimum of the dynamic and the
// static atlas.
//
// Every operation is one IEEE rounding (contraction off, IEEE division and
// square root), so the factor is reproducible bit for bit:
//  - the comparison sampler (LessEqual, linear, clamp to edge) is evaluated
//    as t = uv * S - 0.5, taps floor(t) and floor(t) + 1 clamped to the layer,
//    FP32 bilinear weights of the four 0/1 comparisons;
//  - sin and cos are a Cody-Waite reduction by pi/2 with a three-part
//    constant and Taylor polynomials to degree 9 and 10 (Horner with fused
//    multiply-adds), about 1 ulp for |x| < 6000; the Vogel angle
//    i * GOLDEN_ANGLE + theta goes through the angle-addition identity from
//    sin/cos of theta (once per pixel) and of f32(i) * GOLDEN_ANGLE (a table);
//  - a projection with a NaN coordinate leaves the layer (factor 1), and so
//    does a layer index past the atlas.
#pragma once

#include "hl_internal.hpp"

namespace hl {

__device__ __forceinline__ void sh_sincos(float x, float& s, float& c) {
#pragma clang fp contract(off)
    const float k = __builtin_rintf(x * 0.63661977f);  // x * 2/pi
    float r = x - k * 1.5703125f;                     // pi/2 = 1.5703125 + 4059 * 2^-23 + P3
    r = r - k * 4.8387050628662109375e-4f;
    r = r - k * -4.3711390e-8f;
    const float r2 = r * r;
    float ps = __builtin_fmaf(r2, 2.7557319e-6f, -1.9841270e-4f);
    ps = __builtin_fmaf(r2, ps, 8.3333333e-3f);
    ps = __builtin_fmaf(r2, ps, -1.6666667e-1f);
    const float sr = __builtin_fmaf(r * r2, ps, r);
    float pc = __builtin_fmaf(r2, -2.7557319e-7f, 2.4801587e-5f);
    pc = __builtin_fmaf(r2, pc, -1.3888889e-3f);
    pc = __builtin_fmaf(r2, pc, 4.1666668e-2f);
    pc = __builtin_fmaf(r2, pc, -0.5f);
    const float cr = __builtin_fmaf(r2, pc, 1.0f);
    const int q = (int)k & 3;
    s = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
    c = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
}

__device__ __forceinline__ float sh_fract(float x) {
#pragma clang fp contract(off)
    return x - floorf(x);
}

// hash22, :247-251
__device__ __forceinline__ float sh_hash22(float px, float py) {
#pragma clang fp contract(off)
    const float x = sh_fract(px * 0.1031f), y = sh_fract(py * 0.1031f), z = x;
    float d = x * (y + 33.33f);
    d = d + y * (z + 33.33f);
    d = d + z * (x + 33.33f);
    return sh_fract((x + y) * d);
}

// the shader's min/max/clamp on values that are not NaN
__device__ __forceinline__ float sh_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float sh_max(float a, float b) { return a > b ? a : b; }

// Two adjacent texels of a row: one 8-byte load from a 4-byte aligned address
// (dword alignment is enough for multi-dword global loads).
struct __attribute__((aligned(4))) sh_f2 {
    float x, y;
};

// One tap of the PCF: textureSampleCompareLevel with the LessEqual / linear /
// clamp-to-edge sampler on the dynamic layer dl and the static layer sl
// (size x size, size >= 2) at the same coordinate, min of the two (:314-326).
// Each row's two taps come from one load of the pair starting at
// clamp(x0, 0, size - 2): the left tap is its second texel when
// x0 >= size - 1, the right tap its first when x0 < 0 (clamp to edge); a NaN
// coordinate takes texel 0 for both.  c * w with c in {0, 1} and w >= 0 is
// w or +0 exactly, so the weights are selected, not multiplied.
__device__ __forceinline__ float sh_compare2(const float* dl, const float* sl, uint32_t size, float fs, float u,
                                             float v, float ref) {
#pragma clang fp contract(off)
    const float tx = u * fs - 0.5f, ty = v * fs - 0.5f;
    const float x0f = floorf(tx), y0f = floorf(ty);
    const float fx = tx - x0f, fy = ty - y0f;
    const float hi = (float)(size - 1u);
    // fminf/fmaxf also map a NaN coordinate into the layer
    const uint32_t xs = (uint32_t)fminf(fmaxf(x0f, 0.0f), hi - 1.0f);
    const uint32_t y0 = (uint32_t)fminf(fmaxf(y0f, 0.0f), hi), y1 = (uint32_t)fminf(fmaxf(y0f + 1.0f, 0.0f), hi);
    const uint32_t o0 = y0 * size + xs, o1 = y1 * size + xs;  // < 2^28
    const sh_f2 a0 = *reinterpret_cast<const sh_f2*>(dl + o0), a1 = *reinterpret_cast<const sh_f2*>(dl + o1);
    const sh_f2 b0 = *reinterpret_cast<const sh_f2*>(sl + o0), b1 = *reinterpret_cast<const sh_f2*>(sl + o1);
    const bool left_hi = x0f >= hi, right_lo = !(x0f >= 0.0f);
    const float wl = 1.0f - fx, wt = 1.0f - fy;
    const auto lerp = [&](sh_f2 r0, sh_f2 r1) {
        const float d00 = left_hi ? r0.y : r0.x, d10 = right_lo ? r0.x : r0.y;
        const float d01 = left_hi ? r1.y : r1.x, d11 = right_lo ? r1.x : r1.y;
        const float top = (ref <= d00 ? wl : 0.0f) + (ref <= d10 ? fx : 0.0f);
        const float bot = (ref <= d01 ? wl : 0.0f) + (ref <= d11 ? fx : 0.0f);
        return top * wt + bot * fy;
    };
    return sh_min(lerp(a0, a1), lerp(b0, b1));
}

// light clip, NDC and the shadow uv (:281-290); false outside the layer
__device__ __forceinline__ bool sh_project(const ShadowView& s, uint32_t layer, float wx, float wy, float wz,
                                           float& u, float& v, float& z) {
#pragma clang fp contract(off)
    const hl_gptr<const float> m = hl_g(s.matrices + 16 * (size_t)layer);
    float cl[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float a = m[0 * 4 + r] * wx;
        a = a + m[1 * 4 + r] * wy;
        a = a + m[2 * 4 + r] * wz;
        a = a + m[3 * 4 + r] * 1.0f;
        cl[r] = a;
    }
    if (cl[3] <= 0.0f) return false;
    const float nx = cl[0] / cl[3], ny = cl[1] / cl[3], nz = cl[2] / cl[3];
    u = nx * 0.5f + 0.5f;
    v = -ny * 0.5f + 0.5f;
    z = nz;
    return u >= 0.0f && v >= 0.0f && u <= 1.0f && v <= 1.0f && nz >= 0.0f && nz <= 1.0f;
}

// Vogel disk tables of a context (k_vogel_table): cos and sin of
// f32(i) * GOLDEN_ANGLE for i < kVogelMax, then for every count n <= kVogelMax
// the radii sqrt(f32(i) + 0.5) / sqrt(f32(n)) at [kVogelRadii + n * kVogelMax + i].
constexpr uint32_t kVogelMax = 256;
constexpr uint32_t kVogelRadii = 2 * kVogelMax;
constexpr size_t kVogelFloats = kVogelRadii + (size_t)(kVogelMax + 1) * kVogelMax;
static_assert(kVogelFloats == kVogelTableFloats, "Vogel table size");

// Tap i of count of vogel_disk_sample (:239-244) rotated by theta (sin sth,
// cos cth): cos / sin(i * GOLDEN_ANGLE + theta) by the angle-addition identity.
__device__ __forceinline__ void sh_vogel(const float* tab, uint32_t i, uint32_t count, float sth, float cth,
                                         float& ox, float& oy) {
#pragma clang fp contract(off)
    const float ca = tab[i], sa = tab[kVogelMax + i], r = tab[kVogelRadii + (size_t)count * kVogelMax + i];
    const float c = __builtin_fmaf(ca, cth, -(sa * sth));
    const float sn = __builtin_fmaf(sa, cth, ca * sth);
    ox = c * r;
    oy = sn * r;
}

// PCF of count Vogel taps of the given radius (:306-329)
__device__ __forceinline__ float sh_pcf(const ShadowView& s, uint32_t layer, float u, float v, float ref,
                                        float radius, uint32_t count, float sth, float cth) {
#pragma clang fp contract(off)
    const size_t off = (size_t)layer * s.size * s.size;
    const float* dl = s.atlas + off;
    const float* sl = s.static_atlas + off;
    const float fs = (float)s.size;
    float sum = 0.0f;
    uint32_t i = 0;
    // two taps per step so that their loads are in flight together; the sum
    // keeps the taps' order
    for (; i + 1u < count; i += 2u) {
        float ox0, oy0, ox1, oy1;
        sh_vogel(s.vogel, i, count, sth, cth, ox0, oy0);
        sh_vogel(s.vogel, i + 1u, count, sth, cth, ox1, oy1);
        const float t0 = sh_compare2(dl, sl, s.size, fs, u + ox0 * radius, v + oy0 * radius, ref);
        const float t1 = sh_compare2(dl, sl, s.size, fs, u + ox1 * radius, v + oy1 * radius, ref);
        sum = sum + t0;
        sum = sum + t1;
    }
    if (i < count) {
        float ox, oy;
        sh_vogel(s.vogel, i, count, sth, cth, ox, oy);
        sum = sum + sh_compare2(dl, sl, s.size, fs, u + ox * radius, v + oy * radius, ref);
    }
    return sum / (float)count;
}

// sample_cascade_shadow, :273-330
__device__ __forceinline__ float sh_cascade_pcf(const ShadowView& s, uint32_t layer, uint32_t cascade, float scale,
                                                float wx, float wy, float wz, float sth, float cth) {
#pragma clang fp contract(off)
    if (layer >= s.layers) return 1.0f;
    float u, v, z;
    if (!sh_project(s, layer, wx, wy, wz, u, v, z)) return 1.0f;
    const float radius = (2.0f / (float)s.size) * scale;
    const uint32_t base = s.pcf_count;
    uint32_t n;
    if (cascade == 0u) n = base;
    else if (cascade == 1u) n = base * 3u / 4u > 4u ? base * 3u / 4u : 4u;
    else if (cascade == 2u) n = base / 2u > 4u ? base / 2u : 4u;
    else n = base / 4u > 4u ? base / 4u : 4u;
    return sh_pcf(s, layer, u, v, z, radius, n, sth, cth);
}

// sample_cascade_shadow_pcss with the blocker search and penumbra, :335-450
__device__ __forceinline__ float sh_cascade_pcss(const ShadowView& s, uint32_t layer, uint32_t cascade, float wx,
                                                 float wy, float wz, float sth, float cth) {
#pragma clang fp contract(off)
    if (layer >= s.layers) return 1.0f;
    float u, v, z;
    if (!sh_project(s, layer, wx, wy, wz, u, v, z)) return 1.0f;
    const uint32_t size = s.size;
    const float fs = (float)size;
    const float light_size = s.light_size[cascade], frad = s.filter_radius[cascade];
    const float search = light_size / fs;
    const size_t off = (size_t)layer * size * size;
    const hl_gptr<const float> dl = hl_g(s.atlas + off), sl = hl_g(s.static_atlas + off);
    const uint32_t nb = s.blocker_samples;
    float bsum = 0.0f, bcount = 0.0f;
    for (uint32_t i = 0; i < nb; ++i) {
        float ox, oy;
        sh_vogel(s.vogel, i, nb, sth, cth, ox, oy);
        const float su = u + ox * search, sv = v + oy * search;
        // i32(): toward zero; the bounds test below rejects what lies outside
        const float fx = su * fs, fy = sv * fs;
        if (!(fx > -1.0f && fy > -1.0f && fx < fs && fy < fs)) continue;
        const int px = (int)fx, py = (int)fy;
        const float d = dl[(size_t)py * size + px], st = sl[(size_t)py * size + px];
        const float occ = sh_min(d, st);
        if (occ < z - 0.0001f) {
            bsum = bsum + occ;
            bcount = bcount + 1.0f;
        }
    }
    if (bcount < 0.5f) return 1.0f;
    const float avg = bsum / bcount;
    const float penumbra = (z - avg) / sh_max(avg, 0.001f) * light_size;
    const float radius = sh_min(sh_max(penumbra / fs, frad / fs), frad * 3.0f / fs);
    return sh_pcf(s, layer, u, v, z, radius, s.filter_samples, sth, cth);
}

__device__ __forceinline__ float sh_cascade(const ShadowView& s, uint32_t base, uint32_t c, float wx, float wy,
                                            float wz, float sth, float cth) {
#pragma clang fp contract(off)
    if (s.enable_pcss != 0u && s.light_size[c] > 0.0f) return sh_cascade_pcss(s, base + c, c, wx, wy, wz, sth, cth);
    return sh_cascade_pcf(s, base + c, c, 1.0f + (float)c * 1.5f, wx, wy, wz, sth, cth);
}

// shadow_factor (:452-565) of light record L (position p0.xyz, direction
// p2.xyz, type from p0.w / p2.w) whose first layer is base, at the exact
// world position (wx, wy, wz) with the exact unit normal (nx, ny, nz) and
// fragment coordinate (fx, fy); eye = cameras[0].position_near.xyz.
__device__ __forceinline__ float shadow_factor(const ShadowView& s, const HotLight& L, uint32_t type,
                                               uint32_t base, float wx, float wy, float wz, float nx, float ny,
                                               float nz, float fx, float fy, const float* eye) {
#pragma clang fp contract(off)
    float lx, ly, lz;
    if (type == HL_LIGHT_DIRECTIONAL) {
        lx = -L.p2.x;
        ly = -L.p2.y;
        lz = -L.p2.z;
    } else {
        lx = L.p0.x - wx;
        ly = L.p0.y - wy;
        lz = L.p0.z - wz;
    }
    const float len = sqrtf(hl_x_dot3(lx, ly, lz, lx, ly, lz));
    lx = lx / len;
    ly = ly / len;
    lz = lz / len;
    const float ndl = sh_max(hl_x_dot3(nx, ny, nz, lx, ly, lz), 0.0f);
    const float k = 1.0f - ndl;
    const float bx = wx + nx * 0.01f * k, by = wy + ny * 0.01f * k, bz = wz + nz * 0.01f * k;
    const float theta = sh_hash22(fx, fy) * 6.28318530718f;
    float sth, cth;
    sh_sincos(theta, sth, cth);
    if (type == HL_LIGHT_POINT) {
        const float tx = bx - L.p0.x, ty = by - L.p0.y, tz = bz - L.p0.z;
        const float ax = fabsf(tx), ay = fabsf(ty), az = fabsf(tz);
        uint32_t face;
        if (ax >= ay && ax >= az) face = tx < 0.0f ? 1u : 0u;
        else if (ay >= ax && ay >= az) face = ty < 0.0f ? 3u : 2u;
        else face = tz < 0.0f ? 5u : 4u;
        return sh_cascade_pcf(s, base + face, 0u, 1.0f, bx, by, bz, sth, cth);
    }
    if (type != HL_LIGHT_DIRECTIONAL) return sh_cascade_pcf(s, base, 0u, 1.0f, bx, by, bz, sth, cth);
    // directional: cascades by the camera distance of the unbiased position
    const float dx = wx - eye[0], dy = wy - eye[1], dz = wz - eye[2];
    const float dist = sqrtf(hl_x_dot3(dx, dy, dz, dx, dy, dz));
    const float lo_k = 0.95f, hi_k = 1.05f;  // 1 -+ BLEND_ZONE / 2 (:512), rounded to f32
    uint32_t ca = 3u, cb = 3u;
    float blend = 0.0f;
#pragma unroll
    for (uint32_t j = 0; j < 3u; ++j) {
        const float lo = s.csm[j] * lo_k, hi = s.csm[j] * hi_k;
        if (dist < lo) {
            ca = j;
            break;
        }
        if (dist < hi) {
            ca = j;
            cb = j + 1u;
            blend = hl_x_smoothstep(lo, hi, dist);
            break;
        }
    }
    const float sa = sh_cascade(s, base, ca, bx, by, bz, sth, cth);
    if (blend <= 0.001f) return sa;
    if (cb != ca) {
        const float sb = sh_cascade(s, base, cb, bx, by, bz, sth, cth);
        return sa * (1.0f - blend) + sb * blend;
    }
    return sa;
}

}  // namespace hl
