/*
 * shadow.c -- FP32 restatement of the deferred shader's shadow_factor
 * (crates/passes/3d/helio-pass-deferred-light/shaders/deferred_lighting.wgsl:
 * 234-565): normal-offset bias, cube face / cascade selection, PCF over a
 * Vogel disk and PCSS, each tap the minimum of the dynamic and the static
 * atlas.
 *
 * TEST INFRASTRUCTURE ONLY: used by the tests as the checker of
 * hl_deferred_forward_shadowed, never by the library.  Built with
 * -ffp-contract=off: one IEEE operation per source operation.
 *
 * Fixed here where the shader leaves it to the backend:
 *  - the comparison sampler (LessEqual, linear, clamp to edge;
 *    helio-pass-shadow/src/lib.rs:365-375): texel coordinate t = uv * S - 0.5,
 *    taps floor(t) and floor(t) + 1 clamped to [0, S - 1], FP32 weights
 *    ((c00 (1 - fx) + c10 fx) (1 - fy) + (c01 (1 - fx) + c11 fx) fy), c = 1 when
 *    ref <= texel;
 *  - sin and cos: Cody-Waite reduction by pi/2 (three-part constant) and the
 *    Taylor polynomials to degree 9 / 10, Horner with fused multiply-adds
 *    (or_sincos), accurate to about 1 ulp for |x| < 6000; the Vogel angle
 *    i * GOLDEN_ANGLE + theta by the angle-addition identity from sin/cos of
 *    theta and of f32(i) * GOLDEN_ANGLE (two fused multiply-adds);
 *  - a projection with a NaN coordinate leaves the layer (factor 1).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define SH_NO_SHADOW 0xFFFFFFFFu
#define SH_MAX_SHADOW_LIGHTS 42u

static float sh_fract(float x) { return x - floorf(x); }
static float sh_max(float a, float b) { return a > b ? a : b; }
static float sh_min(float a, float b) { return a < b ? a : b; }
static float sh_clamp(float x, float lo, float hi) { return sh_min(sh_max(x, lo), hi); }

/* pi/2 = P1 + P2 + P3; k * P1 is exact for |k| < 2^16 and k * P2 for |k| < 2^12 */
static const float SH_P1 = 1.5703125f;
static const float SH_P2 = 4.8387050628662109375e-4f; /* 4059 * 2^-23 */
static const float SH_P3 = -4.3711390e-8f;

void or_sincos(float x, float* s, float* c) {
    const float k = rintf(x * 0.63661977f); /* x * 2/pi */
    float r = x - k * SH_P1;
    r = r - k * SH_P2;
    r = r - k * SH_P3;
    const float r2 = r * r;
    float ps = fmaf(r2, 2.7557319e-6f, -1.9841270e-4f); /* 1/9!, -1/7! */
    ps = fmaf(r2, ps, 8.3333333e-3f);                   /* 1/5! */
    ps = fmaf(r2, ps, -1.6666667e-1f);                  /* -1/3! */
    const float sr = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.7557319e-7f, 2.4801587e-5f); /* -1/10!, 1/8! */
    pc = fmaf(r2, pc, -1.3888889e-3f);                  /* -1/6! */
    pc = fmaf(r2, pc, 4.1666668e-2f);                   /* 1/4! */
    pc = fmaf(r2, pc, -0.5f);                           /* -1/2! */
    const float cr = fmaf(r2, pc, 1.0f);
    switch ((int)k & 3) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}

/* hash22, :247-251 */
float or_hash22(float px, float py) {
    const float x = sh_fract(px * 0.1031f), y = sh_fract(py * 0.1031f), z = sh_fract(px * 0.1031f);
    float d = x * (y + 33.33f);
    d = d + y * (z + 33.33f);
    d = d + z * (x + 33.33f);
    return sh_fract((x + y) * d);
}

/* vogel_disk_sample, :239-244, given sin and cos of the rotation theta:
 * cos / sin(i * GOLDEN_ANGLE + theta) by the angle-addition identity from
 * those of a = f32(i) * GOLDEN_ANGLE (a table per sample index in the
 * library), r = sqrt(f32(i) + 0.5) / sqrt(f32(n)). */
static void sh_vogel(uint32_t i, uint32_t n, float sth, float cth, float out[2]) {
    const float r = sqrtf((float)i + 0.5f) / sqrtf((float)n);
    float sa, ca;
    or_sincos((float)i * 2.39996323f, &sa, &ca);
    const float c = fmaf(ca, cth, -(sa * sth));
    const float s = fmaf(sa, cth, ca * sth);
    out[0] = c * r;
    out[1] = s * r;
}

void or_vogel(uint32_t i, uint32_t n, float theta, float out[2]) {
    float sth, cth;
    or_sincos(theta, &sth, &cth);
    sh_vogel(i, n, sth, cth, out);
}

/* point_light_face, :253-263 */
uint32_t or_point_light_face(const float d[3]) {
    const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    if (ax >= ay && ax >= az) return d[0] < 0.0f ? 1u : 0u;
    if (ay >= ax && ay >= az) return d[1] < 0.0f ? 3u : 2u;
    return d[2] < 0.0f ? 5u : 4u;
}

static const float* sh_layer(const float* atlas, uint32_t size, uint32_t layer) {
    return atlas + (size_t)layer * size * size;
}

/* textureSampleCompareLevel with the LessEqual / linear / clamp-to-edge sampler
 * (size >= 2).  A NaN coordinate takes texel 0 for both columns / rows. */
float or_sample_compare(const float* atlas, uint32_t size, uint32_t layer, float u, float v, float ref) {
    const float* t = sh_layer(atlas, size, layer);
    const float S = (float)size;
    const float tx = u * S - 0.5f, ty = v * S - 0.5f;
    const float x0f = floorf(tx), y0f = floorf(ty);
    const float fx = tx - x0f, fy = ty - y0f;
    const float hi = (float)(size - 1u);
    const uint32_t x0 = (uint32_t)sh_clamp(x0f, 0.0f, hi);
    const uint32_t x1 = x0f >= 0.0f ? (uint32_t)sh_clamp(x0f + 1.0f, 0.0f, hi) : 0u;
    const uint32_t y0 = (uint32_t)sh_clamp(y0f, 0.0f, hi), y1 = (uint32_t)sh_clamp(y0f + 1.0f, 0.0f, hi);
    const float c00 = ref <= t[(size_t)y0 * size + x0] ? 1.0f : 0.0f;
    const float c10 = ref <= t[(size_t)y0 * size + x1] ? 1.0f : 0.0f;
    const float c01 = ref <= t[(size_t)y1 * size + x0] ? 1.0f : 0.0f;
    const float c11 = ref <= t[(size_t)y1 * size + x1] ? 1.0f : 0.0f;
    const float top = c00 * (1.0f - fx) + c10 * fx;
    const float bot = c01 * (1.0f - fx) + c11 * fx;
    return top * (1.0f - fy) + bot * fy;
}

/* light clip, NDC and the shadow uv (:281-290); 0 when outside the layer */
static int sh_project(const hl_shadow_maps* sh, uint32_t layer, const float wp[3], float* u, float* v,
                      float* z) {
    const float* m = sh->matrices_d + 16 * (size_t)layer;
    float cl[4];
    for (int r = 0; r < 4; ++r) {
        float s = m[0 * 4 + r] * wp[0];
        s = s + m[1 * 4 + r] * wp[1];
        s = s + m[2 * 4 + r] * wp[2];
        s = s + m[3 * 4 + r] * 1.0f;
        cl[r] = s;
    }
    if (cl[3] <= 0.0f) return 0;
    const float nx = cl[0] / cl[3], ny = cl[1] / cl[3], nz = cl[2] / cl[3];
    *u = nx * 0.5f + 0.5f;
    *v = -ny * 0.5f + 0.5f;
    *z = nz;
    /* written so that a NaN coordinate also leaves the layer (the shader would
     * carry the NaN into the taps) */
    return *u >= 0.0f && *v >= 0.0f && *u <= 1.0f && *v <= 1.0f && nz >= 0.0f && nz <= 1.0f;
}

static float sh_pcf(const hl_shadow_maps* sh, uint32_t layer, float u, float v, float ref, float radius,
                    uint32_t count, float theta) {
    float sum = 0.0f, sth, cth;
    or_sincos(theta, &sth, &cth);
    for (uint32_t i = 0; i < count; ++i) {
        float o[2];
        sh_vogel(i, count, sth, cth, o);
        const float su = u + o[0] * radius, sv = v + o[1] * radius;
        const float dyn = or_sample_compare(sh->atlas_d, sh->atlas_size, layer, su, sv, ref);
        const float sta = or_sample_compare(sh->static_atlas_d, sh->atlas_size, layer, su, sv, ref);
        sum = sum + sh_min(dyn, sta);
    }
    return sum / (float)count;
}

/* sample_cascade_shadow, :273-330 */
float or_sample_cascade_shadow(const hl_shadow_maps* sh, uint32_t layer, uint32_t cascade, float scale,
                               const float wp[3], float fx, float fy) {
    if (layer >= sh->num_layers) return 1.0f;
    float u, v, z;
    if (!sh_project(sh, layer, wp, &u, &v, &z)) return 1.0f;
    const float S = (float)sh->atlas_size;
    const float radius = (2.0f / S) * scale;
    const float theta = or_hash22(fx, fy) * 6.28318530718f;
    const uint32_t base = sh->config.pcf_sample_count;
    uint32_t n;
    switch (cascade) {
        case 0: n = base; break;
        case 1: n = base * 3u / 4u > 4u ? base * 3u / 4u : 4u; break;
        case 2: n = base / 2u > 4u ? base / 2u : 4u; break;
        default: n = base / 4u > 4u ? base / 4u : 4u; break;
    }
    return sh_pcf(sh, layer, u, v, z, radius, n, theta);
}

/* sample_cascade_shadow_pcss with pcss_blocker_search and pcss_penumbra_size, :335-450 */
float or_sample_cascade_shadow_pcss(const hl_shadow_maps* sh, uint32_t layer, uint32_t cascade, const float wp[3],
                                    float fx, float fy) {
    if (layer >= sh->num_layers) return 1.0f;
    const CascadeConfig* cc = &sh->config.cascades[cascade];
    float u, v, z;
    if (!sh_project(sh, layer, wp, &u, &v, &z)) return 1.0f;
    const uint32_t size = sh->atlas_size;
    const float S = (float)size;
    const float theta = or_hash22(fx, fy) * 6.28318530718f;
    const float search = cc->pcss_light_size / S;
    const uint32_t nb = sh->config.pcss_blocker_samples;
    float bsum = 0.0f, bcount = 0.0f, sth, cth;
    or_sincos(theta, &sth, &cth);
    const float* dl = sh_layer(sh->atlas_d, size, layer);
    const float* sl = sh_layer(sh->static_atlas_d, size, layer);
    for (uint32_t i = 0; i < nb; ++i) {
        float o[2];
        sh_vogel(i, nb, sth, cth, o);
        const float su = u + o[0] * search, sv = v + o[1] * search;
        const int px = (int)(su * S), py = (int)(sv * S); /* i32(): toward zero */
        if (px < 0 || py < 0 || px >= (int)size || py >= (int)size) continue;
        const float d = dl[(size_t)py * size + px], s = sl[(size_t)py * size + px];
        const float occ = sh_min(d, s);
        if (occ < z - 0.0001f) {
            bsum = bsum + occ;
            bcount = bcount + 1.0f;
        }
    }
    if (bcount < 0.5f) return 1.0f;
    const float avg = bsum / bcount;
    const float penumbra = (z - avg) / sh_max(avg, 0.001f) * cc->pcss_light_size;
    const float radius = sh_clamp(penumbra / S, cc->filter_radius / S, cc->filter_radius * 3.0f / S);
    return sh_pcf(sh, layer, u, v, z, radius, sh->config.pcss_filter_samples, theta);
}

static float sh_smoothstep(float e0, float e1, float x) {
    const float t = sh_clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

static float sh_cascade(const hl_shadow_maps* sh, uint32_t base, uint32_t c, const float wp[3], float fx,
                        float fy) {
    if (sh->config.enable_pcss != 0u && sh->config.cascades[c].pcss_light_size > 0.0f)
        return or_sample_cascade_shadow_pcss(sh, base + c, c, wp, fx, fy);
    const float scale = 1.0f + (float)c * 1.5f;
    return or_sample_cascade_shadow(sh, base + c, c, scale, wp, fx, fy);
}

/* shadow_factor, :452-565, for lights[light_idx] at world_pos wp with unit
 * normal N, fragment coordinate (fx, fy); cam_pos = cameras[0].position_near. */
float or_shadow_factor(const hl_shadow_maps* sh, const GpuLight* lights, uint32_t light_idx, const float cam_pos[3],
                       const float csm_splits[4], const float wp[3], const float N[3], float fx, float fy) {
    if (!sh || light_idx >= SH_MAX_SHADOW_LIGHTS) return 1.0f;
    const GpuLight* l = &lights[light_idx];
    if (l->shadow_index == SH_NO_SHADOW) return 1.0f;
    float ld[3];
    if (l->light_type == 0u) {
        ld[0] = -l->direction_outer[0];
        ld[1] = -l->direction_outer[1];
        ld[2] = -l->direction_outer[2];
    } else {
        ld[0] = l->position_range[0] - wp[0];
        ld[1] = l->position_range[1] - wp[1];
        ld[2] = l->position_range[2] - wp[2];
    }
    float len = ld[0] * ld[0];
    len = len + ld[1] * ld[1];
    len = len + ld[2] * ld[2];
    len = sqrtf(len);
    ld[0] = ld[0] / len;
    ld[1] = ld[1] / len;
    ld[2] = ld[2] / len;
    float ndl = N[0] * ld[0];
    ndl = ndl + N[1] * ld[1];
    ndl = ndl + N[2] * ld[2];
    ndl = sh_max(ndl, 0.0f);
    float bp[3];
    for (int k = 0; k < 3; ++k) bp[k] = wp[k] + N[k] * 0.01f * (1.0f - ndl);
    const uint32_t base = l->shadow_index;
    if (base >= sh->num_layers) return 1.0f; /* every layer of the light lies past the atlas */
    if (l->light_type == 1u) {
        const float tf[3] = {bp[0] - l->position_range[0], bp[1] - l->position_range[1], bp[2] - l->position_range[2]};
        return or_sample_cascade_shadow(sh, base + or_point_light_face(tf), 0u, 1.0f, bp, fx, fy);
    }
    if (l->light_type != 0u) return or_sample_cascade_shadow(sh, base, 0u, 1.0f, bp, fx, fy);
    /* directional: cascades by camera distance of the unbiased position, :508-558 */
    const float dv[3] = {wp[0] - cam_pos[0], wp[1] - cam_pos[1], wp[2] - cam_pos[2]};
    float dist = dv[0] * dv[0];
    dist = dist + dv[1] * dv[1];
    dist = dist + dv[2] * dv[2];
    dist = sqrtf(dist);
    /* 1 -+ BLEND_ZONE / 2 of :512, abstract constants rounded to f32 */
    const float lo_k = 0.95f, hi_k = 1.05f;
    uint32_t ca = 3u, cb = 3u;
    float blend = 0.0f;
    if (dist < csm_splits[0] * lo_k) {
        ca = 0u;
    } else if (dist < csm_splits[0] * hi_k) {
        ca = 0u;
        cb = 1u;
        blend = sh_smoothstep(csm_splits[0] * lo_k, csm_splits[0] * hi_k, dist);
    } else if (dist < csm_splits[1] * lo_k) {
        ca = 1u;
    } else if (dist < csm_splits[1] * hi_k) {
        ca = 1u;
        cb = 2u;
        blend = sh_smoothstep(csm_splits[1] * lo_k, csm_splits[1] * hi_k, dist);
    } else if (dist < csm_splits[2] * lo_k) {
        ca = 2u;
    } else if (dist < csm_splits[2] * hi_k) {
        ca = 2u;
        cb = 3u;
        blend = sh_smoothstep(csm_splits[2] * lo_k, csm_splits[2] * hi_k, dist);
    }
    const float sa = sh_cascade(sh, base, ca, bp, fx, fy);
    if (blend <= 0.001f) return sa;
    if (cb != ca) {
        const float sb = sh_cascade(sh, base, cb, bp, fx, fy);
        return sa * (1.0f - blend) + sb * blend;
    }
    return sa;
}
