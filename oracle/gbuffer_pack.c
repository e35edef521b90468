/*
 * gbuffer_pack.c -- CPU oracle of the G-buffer producer's format contract.
 *
 * TEST INFRASTRUCTURE ONLY: used by the tests as the checker of
 * hl_gbuffer_pack, never by the library.
 *
 * Restates what the G-buffer pass's fragment shader writes
 * (crates/passes/3d/helio-pass-gbuffer/shaders/gbuffer.wgsl:10-15, 480-493)
 * and its metallic-workflow F0 (gbuffer.wgsl:324-328), with the texel
 * conversions this library fixes (the reference leaves them to the graphics
 * driver): unorm8 = round-to-nearest-even of clamp(v, 0, 1) * 255, NaN -> 0;
 * fp16 = IEEE round-to-nearest-even, written out bit by bit below.
 * Built with -ffp-contract=off: one IEEE operation per source operation.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

/* IEEE binary32 -> binary16, round to nearest, ties to even. */
uint16_t or_float_to_half(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const uint32_t e = (x >> 23) & 0xffu, m = x & 0x7fffffu;
    if (e == 0xffu) return (uint16_t)(sign | (m ? 0x7e00u | (m >> 13) : 0x7c00u));  /* NaN stays NaN */
    if (e >= 143u) return (uint16_t)(sign | 0x7c00u);                            /* >= 2^16: inf */
    if (e >= 113u) {                                                             /* half normal */
        uint32_t h = ((e - 112u) << 10) | (m >> 13);
        const uint32_t rem = m & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;                   /* carry may reach inf */
        return (uint16_t)(sign | h);
    }
    /* half subnormal: value / 2^-24 = (m | 2^23) * 2^(e - 126) */
    if (e == 0u) return sign;                                                    /* float subnormal: < 2^-126 */
    const uint32_t shift = 126u - e;                                             /* >= 14 */
    if (shift >= 25u) return sign;                                               /* < 2^-25 (or exactly): 0 */
    const uint32_t mant = m | 0x800000u;
    uint32_t k = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (k & 1u))) ++k;                            /* 1024 = min normal */
    return (uint16_t)(sign | k);
}

uint32_t or_unorm8_encode(float v) {
    float c = v > 0.0f ? v : 0.0f; /* NaN -> 0 */
    c = c < 1.0f ? c : 1.0f;
    return (uint32_t)rintf(c * 255.0f);
}

/* gbuffer.wgsl:324-328, mix(a, b, t) = a * (1 - t) + b * t */
float or_f0_metallic(float albedo, float metallic) {
    const float m = 0.04f * (1.0f - metallic) + albedo * metallic;
    const float c = m > 0.0f ? m : 0.0f;
    return c < 0.999f ? c : 0.999f;
}

static void store(const hl_texture* t, uint32_t x, uint32_t y, const float v[4]) {
    uint8_t* row = (uint8_t*)t->data_d + (size_t)y * t->pitch_bytes;
    switch (t->format) {
        case HL_FMT_RGBA8_UNORM: {
            const uint32_t w = or_unorm8_encode(v[0]) | (or_unorm8_encode(v[1]) << 8) |
                               (or_unorm8_encode(v[2]) << 16) | (or_unorm8_encode(v[3]) << 24);
            memcpy(row + (size_t)x * 4, &w, 4);
            break;
        }
        case HL_FMT_RGBA16_FLOAT: {
            uint16_t h[4];
            for (int k = 0; k < 4; ++k) h[k] = or_float_to_half(v[k]);
            memcpy(row + (size_t)x * 8, h, 8);
            break;
        }
        default:
            memcpy(row + (size_t)x * 16, v, 16);
            break;
    }
}

static const float* plane_at(const float* p, uint32_t pitch, uint32_t x, uint32_t y) {
    return (const float*)((const uint8_t*)p + (size_t)y * pitch) + 4 * (size_t)x;
}

void or_gbuffer_pack(const hl_surface* s, const hl_gbuffer* g) {
    const uint16_t lm = or_float_to_half(-1.0f);
    for (uint32_t y = 0; y < s->height; ++y)
        for (uint32_t x = 0; x < s->width; ++x) {
            const float* alb = plane_at(s->albedo_d, s->pitch_bytes, x, y);
            const float* n = plane_at(s->normal_d, s->pitch_bytes, x, y);
            const float* orm = plane_at(s->orm_d, s->pitch_bytes, x, y);
            const float* em = plane_at(s->emissive_d, s->pitch_bytes, x, y);
            float f0[3];
            if (s->specular_f0_d) {
                const float* f = plane_at(s->specular_f0_d, s->pitch_bytes, x, y);
                f0[0] = f[0];
                f0[1] = f[1];
                f0[2] = f[2];
            } else {
                for (int k = 0; k < 3; ++k) f0[k] = or_f0_metallic(alb[k], orm[2]);
            }
            const float va[4] = {alb[0], alb[1], alb[2], alb[3]};
            const float vn[4] = {n[0], n[1], n[2], f0[0]};
            const float vo[4] = {orm[0], orm[1], orm[2], f0[1]};
            const float ve[4] = {em[0], em[1], em[2], f0[2]};
            store(&g->albedo, x, y, va);
            store(&g->normal, x, y, vn);
            store(&g->orm, x, y, vo);
            store(&g->emissive, x, y, ve);
            if (g->lightmap_uv.data_d) {
                const uint16_t pair[2] = {lm, lm};
                memcpy((uint8_t*)g->lightmap_uv.data_d + (size_t)y * g->lightmap_uv.pitch_bytes + (size_t)x * 4,
                       pair, 4);
            }
        }
}
