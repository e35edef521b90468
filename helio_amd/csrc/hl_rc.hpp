// hl_rc.hpp -- the deferred shader's radiance-cascade indirect diffuse
// (crates/passes/3d/helio-pass-deferred-light/shaders/deferred_lighting.wgsl:
// 842-904, sample_rc_irradiance) for the shading kernel.  rc_cascade0 is read
// with textureLoad (no filtering); the corner sums are fused multiply-adds
// (WGSL leaves contraction to the backend; the restatement fuses the same
// ones) and every other operation is one IEEE rounding, so the irradiance
// agrees with the restatement bit for bit.
#pragma once

#include "hl_internal.hpp"

namespace hl {

constexpr uint32_t kRcProbeDim = 16u;  // RC_PROBE_DIM, :824
constexpr uint32_t kRcDirDim = 4u;     // RC_DIR_DIM, :825

// rc_oct_decode (:827-840) of the 16 probe directions ((dx, dy) + 0.5) / 4,
// index dx * 4 + dy, as the shader's FP32 operations round them (exact
// values; they fold into the instructions as literals).
__device__ constexpr float kRcDir[16][3] = {
    {-0x1.a20bd60000000p-2f, -0x1.a20bd60000000p-1f, -0x1.a20bd60000000p-2f},
    {-0x1.e5b9d00000000p-1f, 0x0.0p+0f, -0x1.43d1360000000p-2f},
    {-0x1.e5b9d00000000p-1f, 0x0.0p+0f, 0x1.43d1360000000p-2f},
    {-0x1.a20bd60000000p-2f, -0x1.a20bd60000000p-1f, 0x1.a20bd60000000p-2f},
    {-0x1.43d1360000000p-2f, 0x0.0p+0f, -0x1.e5b9d00000000p-1f},
    {-0x1.a20bd60000000p-2f, 0x1.a20bd60000000p-1f, -0x1.a20bd60000000p-2f},
    {-0x1.a20bd60000000p-2f, 0x1.a20bd60000000p-1f, 0x1.a20bd60000000p-2f},
    {-0x1.43d1360000000p-2f, 0x0.0p+0f, 0x1.e5b9d00000000p-1f},
    {0x1.43d1360000000p-2f, 0x0.0p+0f, -0x1.e5b9d00000000p-1f},
    {0x1.a20bd60000000p-2f, 0x1.a20bd60000000p-1f, -0x1.a20bd60000000p-2f},
    {0x1.a20bd60000000p-2f, 0x1.a20bd60000000p-1f, 0x1.a20bd60000000p-2f},
    {0x1.43d1360000000p-2f, 0x0.0p+0f, 0x1.e5b9d00000000p-1f},
    {0x1.a20bd60000000p-2f, -0x1.a20bd60000000p-1f, -0x1.a20bd60000000p-2f},
    {0x1.e5b9d00000000p-1f, 0x0.0p+0f, -0x1.43d1360000000p-2f},
    {0x1.e5b9d00000000p-1f, 0x0.0p+0f, 0x1.43d1360000000p-2f},
    {0x1.a20bd60000000p-2f, -0x1.a20bd60000000p-1f, 0x1.a20bd60000000p-2f},
};

__device__ __forceinline__ float rc_smoothstep(float e0, float e1, float x) {
#pragma clang fp contract(off)
    const float t = fminf(fmaxf((x - e0) / (e1 - e0), 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// rgb of texel (x, y) of the probe atlas (RGBA16F or RGBA32F)
__device__ __forceinline__ float3 rc_texel(const Tex& t, uint32_t fmt, uint32_t x, uint32_t y) {
    const uint8_t* row = t.data + (size_t)y * t.pitch;
    if (fmt == HL_FMT_RGBA16_FLOAT) {
        const hl_u32x2 w = *hl_g(reinterpret_cast<const hl_u32x2*>(row) + x);
        return make_float3(hl_half((uint16_t)(w.x & 0xffffu)), hl_half((uint16_t)(w.x >> 16)),
                           hl_half((uint16_t)(w.y & 0xffffu)));
    }
    const float4 v = hl_ld4(reinterpret_cast<const float4*>(row) + x);
    return make_float3(v.x, v.y, v.z);
}

// sample_rc_irradiance (:866-904) at world position (wx, wy, wz) with unit
// normal (nx, ny, nz); the caller checks has_rc_gi.  sample_rc_specular
// (:929-983) is the same body along its lookup direction.  A: any launch
// record with the atlas (rc, rc_fmt) and its box (rc_min, rc_size).
template <class A>
__device__ __forceinline__ float3 rc_irradiance(const A& a, float wx, float wy, float wz, float nx, float ny,
                                                float nz) {
#pragma clang fp contract(off)
    const float sx = a.rc_size[0], sy = a.rc_size[1], sz = a.rc_size[2];
    if (sx <= 0.0f || sy <= 0.0f || sz <= 0.0f) return make_float3(0.0f, 0.0f, 0.0f);
    const float dx = wx - a.rc_min[0], dy = wy - a.rc_min[1], dz = wz - a.rc_min[2];
    const float tx = dx / sx, ty = dy / sy, tz = dz / sz;
    const float vw = rc_smoothstep(0.0f, 0.05f, tx) * rc_smoothstep(1.0f, 0.95f, tx) *
                     (rc_smoothstep(0.0f, 0.05f, ty) * rc_smoothstep(1.0f, 0.95f, ty)) *
                     (rc_smoothstep(0.0f, 0.05f, tz) * rc_smoothstep(1.0f, 0.95f, tz));
    if (vw <= 0.0f) return make_float3(0.0f, 0.0f, 0.0f);
    float cw[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) cw[k] = fmaxf(0.0f, hl_x_dot3(nx, ny, nz, kRcDir[k][0], kRcDir[k][1], kRcDir[k][2]));
    const float hi = (float)kRcProbeDim - 1.0f;
    const float pfx = fminf(fmaxf(dx / (sx / (float)kRcProbeDim) - 0.5f, 0.0f), hi);
    const float pfy = fminf(fmaxf(dy / (sy / (float)kRcProbeDim) - 0.5f, 0.0f), hi);
    const float pfz = fminf(fmaxf(dz / (sz / (float)kRcProbeDim) - 0.5f, 0.0f), hi);
    const uint32_t px = (uint32_t)pfx, py = (uint32_t)pfy, pz = (uint32_t)pfz;
    const float fx = pfx - floorf(pfx), fy = pfy - floorf(pfy), fz = pfz - floorf(pfz);
    const auto mix = [](float p, float q, float t) {
#pragma clang fp contract(off)
        return p * (1.0f - t) + q * t;
    };
    // rc_corner_irradiance_precomp (:842-864) of corner (x, y, z) offsets; the
    // sum runs over direction k = dx * 4 + dy in the shader's order
    const auto corner = [&](uint32_t ox, uint32_t oy, uint32_t oz) {
#pragma clang fp contract(off)
        const uint32_t dim = kRcProbeDim - 1u;
        const uint32_t cx = min(px + ox, dim), cy = min(py + oy, dim), cz = min(pz + oz, dim);
        const uint32_t row0 = (cy * kRcProbeDim + cz) * kRcDirDim;
        float ir = 0.0f, ig = 0.0f, ib = 0.0f, ws = 0.0f;
        const auto add = [&](uint32_t k, float r, float g, float b) {
            const float w = cw[k];
            if (w > 0.001f) {
                ir = __builtin_fmaf(r, w, ir);
                ig = __builtin_fmaf(g, w, ig);
                ib = __builtin_fmaf(b, w, ib);
                ws = ws + w;
            }
        };
        if (a.rc_fmt == HL_FMT_RGBA16_FLOAT) {
            // the block's rows are 4 texels (32 B): two 16-B loads per row,
            // columns dx = 2h, 2h + 1 in the load of half h
#pragma unroll
            for (uint32_t h = 0; h < 2u; ++h) {
                hl_u32x4 q[4];
#pragma unroll
                for (uint32_t dy = 0; dy < 4u; ++dy)
                    q[dy] = *hl_g(reinterpret_cast<const hl_u32x4*>(a.rc.data + (size_t)(row0 + dy) * a.rc.pitch) +
                                  (cx * 2u + h));
#pragma unroll
                for (uint32_t dy = 0; dy < 4u; ++dy)
                    add((2u * h) * 4u + dy, hl_half((uint16_t)(q[dy].x & 0xffffu)), hl_half((uint16_t)(q[dy].x >> 16)),
                        hl_half((uint16_t)(q[dy].y & 0xffffu)));
#pragma unroll
                for (uint32_t dy = 0; dy < 4u; ++dy)
                    add((2u * h + 1u) * 4u + dy, hl_half((uint16_t)(q[dy].z & 0xffffu)),
                        hl_half((uint16_t)(q[dy].z >> 16)), hl_half((uint16_t)(q[dy].w & 0xffffu)));
            }
        } else {
#pragma unroll
            for (uint32_t k = 0; k < 16u; ++k) {
                if (cw[k] > 0.001f) {
                    const float3 t = rc_texel(a.rc, a.rc_fmt, cx * kRcDirDim + (k >> 2), row0 + (k & 3u));
                    add(k, t.x, t.y, t.z);
                }
            }
        }
        const float d = fmaxf(ws, 0.001f);
        return make_float3(ir / d, ig / d, ib / d);
    };
    // mix(mix(c000, c001, fz), mix(c010, c011, fz), fy) and the same for x + 1,
    // each z pair folded as soon as both corners are known
    const auto zpair = [&](uint32_t ox, uint32_t oy) {
        const float3 c0 = corner(ox, oy, 0u);
        const float3 c1 = corner(ox, oy, 1u);
        return make_float3(mix(c0.x, c1.x, fz), mix(c0.y, c1.y, fz), mix(c0.z, c1.z, fz));
    };
    const auto ypair = [&](uint32_t ox) {
        const float3 m0 = zpair(ox, 0u);
        const float3 m1 = zpair(ox, 1u);
        return make_float3(mix(m0.x, m1.x, fy), mix(m0.y, m1.y, fy), mix(m0.z, m1.z, fy));
    };
    const float3 x0 = ypair(0u);
    const float3 x1 = ypair(1u);
    return make_float3(mix(x0.x, x1.x, fx) * vw, mix(x0.y, x1.y, fx) * vw, mix(x0.z, x1.z, fx) * vw);
}

}  // namespace hl
