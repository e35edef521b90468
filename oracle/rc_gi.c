/*
 * rc_gi.c -- FP32 restatement of the deferred shader's radiance-cascade
 * indirect diffuse (crates/passes/3d/helio-pass-deferred-light/shaders/
 * deferred_lighting.wgsl:824-904 sample_rc_irradiance, and its combine at
 * :1159-1163, 1189-1205).
 *
 * TEST INFRASTRUCTURE ONLY: used by the tests as the checker of
 * hl_deferred_forward_ex, never by the library.  Built with
 * -ffp-contract=off: one IEEE operation per source operation, except the
 * corner sums, written as fused multiply-adds (fmaf): WGSL leaves contraction
 * to the backend, and the kernel fuses exactly these.
 *
 * rc_cascade0 is read with textureLoad (no filtering), so the restatement is
 * exact: the probe atlas is RC_DIR_DIM * RC_PROBE_DIM = 64 texels wide and
 * RC_PROBE_DIM^2 * RC_DIR_DIM = 1024 rows, RGBA16F or RGBA32F.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define RC_PROBE_DIM 16u
#define RC_DIR_DIM 4u

static float rc_max(float a, float b) { return a > b ? a : b; }
static float rc_min(float a, float b) { return a < b ? a : b; }
static float rc_clamp(float x, float lo, float hi) { return rc_min(rc_max(x, lo), hi); }
static float rc_smoothstep(float e0, float e1, float x) {
    const float t = rc_clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
static float rc_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

static float rc_half(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t bits;
    if (e == 0u) {
        if (m == 0u) {
            bits = s;
        } else { /* subnormal: normalise */
            uint32_t mm = m, ee = 113u;
            while (!(mm & 0x400u)) {
                mm <<= 1;
                --ee;
            }
            bits = s | (ee << 23) | ((mm & 0x3ffu) << 13);
        }
    } else if (e == 31u) {
        bits = s | 0x7f800000u | (m << 13);
    } else {
        bits = s | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

/* rc_oct_decode, :827-840 */
void or_rc_oct_decode(float u, float v, float out[3]) {
    const float fx = u * 2.0f - 1.0f, fy = v * 2.0f - 1.0f;
    const float ax = fabsf(fx), ay = fabsf(fy);
    const float l = ax + ay;
    float n[3];
    if (l > 1.0f) {
        const float sx = fx >= 0.0f ? 1.0f : -1.0f, sz = fy >= 0.0f ? 1.0f : -1.0f;
        n[0] = (1.0f - ay) * sx;
        n[1] = 1.0f - l;
        n[2] = (1.0f - ax) * sz;
    } else {
        n[0] = fx;
        n[1] = 1.0f - l;
        n[2] = fy;
    }
    float d = n[0] * n[0];
    d = d + n[1] * n[1];
    d = d + n[2] * n[2];
    const float len = sqrtf(d);
    out[0] = n[0] / len;
    out[1] = n[1] / len;
    out[2] = n[2] / len;
}

static void rc_texel(const hl_texture* t, uint32_t x, uint32_t y, float rgb[3]) {
    const uint8_t* row = (const uint8_t*)t->data_d + (size_t)y * t->pitch_bytes;
    if (t->format == HL_FMT_RGBA16_FLOAT) {
        const uint16_t* p = (const uint16_t*)row + 4 * (size_t)x;
        for (int k = 0; k < 3; ++k) rgb[k] = rc_half(p[k]);
    } else {
        const float* p = (const float*)row + 4 * (size_t)x;
        for (int k = 0; k < 3; ++k) rgb[k] = p[k];
    }
}

/* rc_corner_irradiance_precomp, :842-864 */
static void rc_corner(const hl_texture* t, uint32_t px, uint32_t py, uint32_t pz, const float cw[16], float out[3]) {
    const uint32_t dim = RC_PROBE_DIM - 1u;
    const uint32_t cpx = px < dim ? px : dim, cpy = py < dim ? py : dim, cpz = pz < dim ? pz : dim;
    float irr[3] = {0.0f, 0.0f, 0.0f}, wsum = 0.0f;
    uint32_t idx = 0;
    for (uint32_t dx = 0; dx < RC_DIR_DIM; ++dx)
        for (uint32_t dy = 0; dy < RC_DIR_DIM; ++dy) {
            const float w = cw[idx++];
            if (w > 0.001f) {
                float c[3];
                rc_texel(t, cpx * RC_DIR_DIM + dx, (cpy * RC_PROBE_DIM + cpz) * RC_DIR_DIM + dy, c);
                for (int k = 0; k < 3; ++k) irr[k] = fmaf(c[k], w, irr[k]); /* contracted, as the kernel */
                wsum = wsum + w;
            }
        }
    const float d = rc_max(wsum, 0.001f);
    for (int k = 0; k < 3; ++k) out[k] = irr[k] / d;
}

/* sample_rc_irradiance, :866-904 (the caller checks has_rc_gi) */
void or_rc_irradiance(const hl_texture* t, const float wmin[3], const float wmax[3], const float wp[3],
                      const float N[3], float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    float size[3], tt[3];
    for (int k = 0; k < 3; ++k) size[k] = wmax[k] - wmin[k];
    if (size[0] <= 0.0f || size[1] <= 0.0f || size[2] <= 0.0f) return;
    float vw = 1.0f;
    for (int k = 0; k < 3; ++k) {
        tt[k] = (wp[k] - wmin[k]) / size[k];
        const float fade = rc_smoothstep(0.0f, 0.05f, tt[k]) * rc_smoothstep(1.0f, 1.0f - 0.05f, tt[k]);
        vw = k == 0 ? fade : vw * fade;
    }
    if (vw <= 0.0f) return;
    float cw[16];
    uint32_t idx = 0;
    for (uint32_t dx = 0; dx < RC_DIR_DIM; ++dx)
        for (uint32_t dy = 0; dy < RC_DIR_DIM; ++dy) {
            float d[3];
            or_rc_oct_decode(((float)dx + 0.5f) / (float)RC_DIR_DIM, ((float)dy + 0.5f) / (float)RC_DIR_DIM, d);
            float c = N[0] * d[0];
            c = c + N[1] * d[1];
            c = c + N[2] * d[2];
            cw[idx++] = rc_max(0.0f, c);
        }
    uint32_t pi[3];
    float fr[3];
    for (int k = 0; k < 3; ++k) {
        const float cell = size[k] / (float)RC_PROBE_DIM;
        const float pf = rc_clamp((wp[k] - wmin[k]) / cell - 0.5f, 0.0f, (float)RC_PROBE_DIM - 1.0f);
        pi[k] = (uint32_t)pf;
        fr[k] = pf - floorf(pf);
    }
    float c[8][3];
    for (uint32_t b = 0; b < 8; ++b) /* b = x << 2 | y << 1 | z */
        rc_corner(t, pi[0] + (b >> 2), pi[1] + ((b >> 1) & 1u), pi[2] + (b & 1u), cw, c[b]);
    for (int k = 0; k < 3; ++k) {
        const float c0 = rc_mix(rc_mix(c[0][k], c[1][k], fr[2]), rc_mix(c[2][k], c[3][k], fr[2]), fr[1]);
        const float c1 = rc_mix(rc_mix(c[4][k], c[5][k], fr[2]), rc_mix(c[6][k], c[7][k], fr[2]), fr[1]);
        out[k] = rc_mix(c0, c1, fr[0]) * vw;
    }
}
