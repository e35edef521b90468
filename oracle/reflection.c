/*
 * reflection.c -- FP32 restatement of the deferred pass's second draw,
 * fs_reflection (crates/passes/3d/helio-pass-deferred-light/shaders/
 * deferred_lighting.wgsl:1266-1341), with sample_reflection_environment
 * (:646-691), capture_weight (:619-639), parallax_sphere / parallax_box
 * (:577-615), sample_rc_specular (:929-983) and env_brdf_approx
 * (crates/libhelio/shaders/pbr_eval.wgsl:70-76).
 *
 * TEST INFRASTRUCTURE ONLY: used by the tests as the checker of
 * hl_deferred_reflection, never by the library.  Built with -ffp-contract=off:
 * one IEEE operation per source operation.
 *
 * The cube sampler (linear, linear mips, clamp to edge;
 * crates/helio-bake/src/data.rs:372-380) is defined here as: the face of the
 * direction's major axis (x on ties with y or z, then y on a tie with z), the
 * face coordinates of the usual cube-map table, bilinear weights in FP32 on
 * the texel coordinate uv * size - 0.5 clamped to the face's edge, and the two
 * mips floor(lod), floor(lod) + 1 mixed by fract(lod) (lod clamped to the mip
 * range).  Hardware samplers quantise the weights and may filter across face
 * edges; both are implementation-defined in WebGPU.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

#define ENV_MAX_LOD 8.0f /* pbr_eval.wgsl:85 */

static float rf_max(float a, float b) { return a > b ? a : b; }
static float rf_min(float a, float b) { return a < b ? a : b; }
/* clamp with NaN mapped to lo (fmaxf/fminf semantics: the non-NaN operand) */
static float rf_clamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
static float rf_dot3(const float a[3], const float b[3]) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}
static void rf_normalize(const float v[3], float o[3]) {
    const float l = sqrtf(rf_dot3(v, v));
    o[0] = v[0] / l;
    o[1] = v[1] / l;
    o[2] = v[2] / l;
}
static float rf_mat_row(const float* m, int row, float x, float y, float z, float w) {
    float s = m[0 * 4 + row] * x;
    s = s + m[1 * 4 + row] * y;
    s = s + m[2 * 4 + row] * z;
    s = s + m[3 * 4 + row] * w;
    return s;
}
static float rf_smoothstep(float e0, float e1, float x) {
    const float t = rf_clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
static float rf_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }

static float rf_half(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t bits;
    if (e == 0u) {
        if (m == 0u) {
            bits = s;
        } else {
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

static uint32_t rf_texel_bytes(uint32_t fmt) {
    return fmt == HL_FMT_RGBA32_FLOAT ? 16u : fmt == HL_FMT_RGBA16_FLOAT ? 8u : 4u;
}

static void rf_decode(const uint8_t* p, uint32_t fmt, float o[4]) {
    if (fmt == HL_FMT_RGBA32_FLOAT) {
        memcpy(o, p, 16);
    } else if (fmt == HL_FMT_RGBA16_FLOAT) {
        uint16_t h[4];
        memcpy(h, p, 8);
        for (int k = 0; k < 4; ++k) o[k] = rf_half(h[k]);
    } else {
        for (int k = 0; k < 4; ++k) o[k] = (float)p[k] / 255.0f;
    }
}

size_t or_cube_bytes(uint32_t face_size, uint32_t mip_levels, uint32_t format) {
    size_t b = 0;
    for (uint32_t m = 0; m < mip_levels; ++m) {
        const size_t r = (face_size >> m) ? (face_size >> m) : 1u;
        b += 6u * r * r * rf_texel_bytes(format);
    }
    return b;
}

/* face selection and face coordinates of direction d */
void or_cube_face(const float d[3], uint32_t* face, float* u, float* v) {
    const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    float ma, sc, tc;
    if (ax >= ay && ax >= az) {
        *face = d[0] >= 0.0f ? 0u : 1u;
        ma = ax;
        sc = d[0] >= 0.0f ? -d[2] : d[2];
        tc = -d[1];
    } else if (ay >= az) {
        *face = d[1] >= 0.0f ? 2u : 3u;
        ma = ay;
        sc = d[0];
        tc = d[1] >= 0.0f ? d[2] : -d[2];
    } else {
        *face = d[2] >= 0.0f ? 4u : 5u;
        ma = az;
        sc = d[2] >= 0.0f ? d[0] : -d[0];
        tc = -d[1];
    }
    *u = (sc / ma + 1.0f) * 0.5f;
    *v = (tc / ma + 1.0f) * 0.5f;
}

/* bilinear sample of one face of one mip, clamped to the face's edge */
static void rf_bilinear(const hl_cube_array* env, const uint8_t* mip_base, uint32_t res, uint32_t face, float u,
                        float v, float o[3]) {
    const uint32_t tb = rf_texel_bytes(env->format);
    const float fs = (float)res;
    const float x = u * fs - 0.5f, y = v * fs - 0.5f;
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const float hi = fs - 1.0f;
    const uint32_t ix0 = (uint32_t)rf_clamp(x0, 0.0f, hi), ix1 = (uint32_t)rf_clamp(x0 + 1.0f, 0.0f, hi);
    const uint32_t iy0 = (uint32_t)rf_clamp(y0, 0.0f, hi), iy1 = (uint32_t)rf_clamp(y0 + 1.0f, 0.0f, hi);
    const uint8_t* fb = mip_base + (size_t)face * res * res * tb;
    float t00[4], t10[4], t01[4], t11[4];
    rf_decode(fb + ((size_t)iy0 * res + ix0) * tb, env->format, t00);
    rf_decode(fb + ((size_t)iy0 * res + ix1) * tb, env->format, t10);
    rf_decode(fb + ((size_t)iy1 * res + ix0) * tb, env->format, t01);
    rf_decode(fb + ((size_t)iy1 * res + ix1) * tb, env->format, t11);
    for (int k = 0; k < 3; ++k) {
        const float top = t00[k] * (1.0f - fx) + t10[k] * fx;
        const float bot = t01[k] * (1.0f - fx) + t11[k] * fx;
        o[k] = top * (1.0f - fy) + bot * fy;
    }
}

/* textureSampleLevel(env_cube, env_sampler, dir, cube, lod).rgb */
void or_cube_sample(const hl_cube_array* env, uint32_t cube, const float dir[3], float lod, float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    if (!env->data_d || env->num_cubes == 0) return;
    if (cube >= env->num_cubes) cube = env->num_cubes - 1u;
    uint32_t face;
    float u, v;
    or_cube_face(dir, &face, &u, &v);
    const float l = rf_clamp(lod, 0.0f, (float)(env->mip_levels - 1u));
    const float lf = floorf(l);
    const uint32_t m0 = (uint32_t)lf, m1 = m0 + 1u < env->mip_levels ? m0 + 1u : m0;
    const float t = l - lf;
    const uint32_t tb = rf_texel_bytes(env->format);
    const uint8_t* cb = (const uint8_t*)env->data_d + or_cube_bytes(env->face_size, env->mip_levels, env->format) * cube;
    size_t off0 = 0;
    for (uint32_t m = 0; m < m0; ++m) {
        const size_t r = (env->face_size >> m) ? (env->face_size >> m) : 1u;
        off0 += 6u * r * r * tb;
    }
    const uint32_t r0 = (env->face_size >> m0) ? (env->face_size >> m0) : 1u;
    const uint32_t r1 = (env->face_size >> m1) ? (env->face_size >> m1) : 1u;
    const size_t off1 = m1 == m0 ? off0 : off0 + 6u * (size_t)r0 * r0 * tb;
    float c0[3], c1[3];
    rf_bilinear(env, cb + off0, r0, face, u, v, c0);
    rf_bilinear(env, cb + off1, r1, face, u, v, c1);
    for (int k = 0; k < 3; ++k) out[k] = rf_mix(c0[k], c1[k], t);
}

/* parallax_sphere (:577-591) */
static void rf_parallax_sphere(const float P[3], const float R[3], const float c[3], float radius, float o[3]) {
    const float oc[3] = {P[0] - c[0], P[1] - c[1], P[2] - c[2]};
    const float b = rf_dot3(oc, R);
    const float cc = rf_dot3(oc, oc) - radius * radius;
    const float disc = b * b - cc;
    if (disc < 0.0f) {
        memcpy(o, R, 12);
        return;
    }
    const float t = -b + sqrtf(disc);
    if (t <= 0.0f) {
        memcpy(o, R, 12);
        return;
    }
    for (int k = 0; k < 3; ++k) o[k] = (P[k] + R[k] * t) - c[k];
}

/* parallax_box (:599-615); max/min of NaN keep the other operand */
static void rf_parallax_box(const float P[3], const float R[3], const GpuReflectionCapture* cap, float o[3]) {
    const float* m = cap->world_to_local;
    float pl[3], rl[3], tv[3];
    for (int r = 0; r < 3; ++r) {
        pl[r] = rf_mat_row(m, r, P[0], P[1], P[2], 1.0f);
        rl[r] = rf_mat_row(m, r, R[0], R[1], R[2], 0.0f);
    }
    for (int k = 0; k < 3; ++k) {
        const float e = cap->extents_transition[k];
        const float inv = 1.0f / rl[k];
        const float t1 = (-e - pl[k]) * inv, t2 = (e - pl[k]) * inv;
        tv[k] = fmaxf(t1, t2);
    }
    const float t = fminf(fminf(tv[0], tv[1]), tv[2]);
    if (t <= 0.0f || !(t < 3.402823e38f)) {
        memcpy(o, R, 12);
        return;
    }
    for (int k = 0; k < 3; ++k) o[k] = (P[k] + R[k] * t) - cap->position_radius[k];
}

void or_capture_parallax(const float P[3], const float R[3], const GpuReflectionCapture* cap, float o[3]) {
    if (cap->shape == HL_CAPTURE_SHAPE_BOX) rf_parallax_box(P, R, cap, o);
    else rf_parallax_sphere(P, R, cap->position_radius, cap->position_radius[3], o);
}

/* capture_weight (:619-639) */
float or_capture_weight(const float P[3], const GpuReflectionCapture* cap) {
    if (cap->shape == HL_CAPTURE_SHAPE_BOX) {
        const float* m = cap->world_to_local;
        float d[3];
        for (int r = 0; r < 3; ++r) d[r] = cap->extents_transition[r] - fabsf(rf_mat_row(m, r, P[0], P[1], P[2], 1.0f));
        if (d[0] < 0.0f || d[1] < 0.0f || d[2] < 0.0f) return 0.0f;
        const float tr = rf_max(cap->extents_transition[3], 1e-4f);
        return rf_clamp(rf_min(rf_min(d[0], d[1]), d[2]) / tr, 0.0f, 1.0f);
    }
    const float radius = cap->position_radius[3];
    const float dv[3] = {P[0] - cap->position_radius[0], P[1] - cap->position_radius[1],
                         P[2] - cap->position_radius[2]};
    const float dist = sqrtf(rf_dot3(dv, dv));
    if (dist > radius) return 0.0f;
    const float fade = rf_max(radius * 0.1f, 1e-4f);
    return rf_clamp((radius - dist) / fade, 0.0f, 1.0f);
}

/* sample_reflection_environment (:646-691) */
void or_reflection_env(const DeferredGlobals* g, const hl_reflection_inputs* in, const float P[3], const float R[3],
                       float lod, float out[3]) {
    float acc[3] = {0.0f, 0.0f, 0.0f}, acc_a = 0.0f;
    out[0] = out[1] = out[2] = 0.0f;
    if (g->enable_reflections == 0u) return;
    const uint32_t count = g->reflection_capture_count < 64u ? g->reflection_capture_count : 64u;
    for (uint32_t i = 0; i < count; ++i) {
        if (acc_a >= 0.999f) break;
        const GpuReflectionCapture* cap = in->captures_d + i;
        if (cap->cubemap_index < 0) continue;
        const float w = or_capture_weight(P, cap);
        if (w <= 0.0f) continue;
        float dir[3], s[3];
        or_capture_parallax(P, R, cap, dir);
        or_cube_sample(&in->env, (uint32_t)cap->cubemap_index, dir, lod, s);
        const float contrib = w * (1.0f - acc_a);
        for (int k = 0; k < 3; ++k) acc[k] = acc[k] + (s[k] * cap->brightness) * contrib;
        acc_a = acc_a + contrib;
    }
    if (acc_a < 0.999f) {
        float sky[3];
        or_cube_sample(&in->env, 0u, R, lod, sky);
        for (int k = 0; k < 3; ++k) acc[k] = acc[k] + sky[k] * (1.0f - acc_a);
    }
    memcpy(out, acc, 12);
}

/* env_brdf_approx (pbr_eval.wgsl:70-76) */
void or_env_brdf_approx(float ndv, float roughness, float out[2]) {
    const float r0 = roughness * -1.0f + 1.0f, r1 = roughness * -0.0275f + 0.0425f;
    const float r2 = roughness * -0.572f + 1.04f, r3 = roughness * 0.022f + -0.04f;
    const float a004 = rf_min(r0 * r0, exp2f(-9.28f * ndv)) * r0 + r1;
    out[0] = -1.04f * a004 + r2;
    out[1] = 1.04f * a004 + r3;
}

static uint32_t rf_debug_zero(uint32_t m) {
    return m == 1u || m == 2u || m == 4u || m == 5u || m == 10u || m == 11u || m == 20u || m == 21u;
}

/* fs_reflection over rows [row0, row1): returns the number of pixels written */
uint64_t or_reflection(const GpuCameraUniforms* cam, const DeferredGlobals* g, const hl_gbuffer* gb,
                       const hl_reflection_inputs* in, float* out, uint32_t out_pitch, uint32_t flags,
                       uint32_t row0, uint32_t row1) {
    const uint32_t W = gb->width, H = gb->height;
    if (row1 > H) row1 = H;
    uint64_t written = 0;
    for (uint32_t y = row0; y < row1; ++y) {
        for (uint32_t x = 0; x < W; ++x) {
            float* o = (float*)((uint8_t*)out + (size_t)y * out_pitch) + 4 * (size_t)x;
            float depth;
            memcpy(&depth, (const uint8_t*)gb->depth.data_d + (size_t)y * gb->depth.pitch_bytes + 4 * (size_t)x, 4);
            if (depth >= 1.0f) continue; /* discard */
            float ssr[4] = {0.0f, 0.0f, 0.0f, 0.0f}, pl[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (in->ssr.data_d)
                rf_decode((const uint8_t*)in->ssr.data_d + (size_t)y * in->ssr.pitch_bytes +
                              (size_t)x * rf_texel_bytes(in->ssr.format),
                          in->ssr.format, ssr);
            if (g->debug_mode == 30u) {
                o[0] = o[1] = o[2] = ssr[3];
                o[3] = 1.0f;
                ++written;
                continue;
            }
            if (g->debug_mode == 31u) {
                o[0] = ssr[0];
                o[1] = ssr[1];
                o[2] = ssr[2];
                o[3] = 1.0f;
                ++written;
                continue;
            }
            if (rf_debug_zero(g->debug_mode)) {
                if (flags & HL_REFLECTION_STORE) {
                    o[0] = o[1] = o[2] = o[3] = 0.0f;
                    ++written;
                }
                continue;
            }
            float nr[4], orm[4], em[4];
            rf_decode((const uint8_t*)gb->normal.data_d + (size_t)y * gb->normal.pitch_bytes +
                          (size_t)x * rf_texel_bytes(gb->normal.format),
                      gb->normal.format, nr);
            rf_decode((const uint8_t*)gb->orm.data_d + (size_t)y * gb->orm.pitch_bytes +
                          (size_t)x * rf_texel_bytes(gb->orm.format),
                      gb->orm.format, orm);
            rf_decode((const uint8_t*)gb->emissive.data_d + (size_t)y * gb->emissive.pitch_bytes +
                          (size_t)x * rf_texel_bytes(gb->emissive.format),
                      gb->emissive.format, em);
            float N[3];
            rf_normalize(nr, N);
            const float rough = orm[1];
            const float F0[3] = {rf_clamp(nr[3], 0.0f, 0.999f), rf_clamp(orm[3], 0.0f, 0.999f),
                                 rf_clamp(em[3], 0.0f, 0.999f)};
            const float ao = orm[0] * 1.0f; /* screen-space AO: the white fallback */
            float P[3];
            or_world_pos(cam, W, H, x, y, depth, P);
            const float e[3] = {cam->position_near[0] - P[0], cam->position_near[1] - P[1],
                                cam->position_near[2] - P[2]};
            float V[3];
            rf_normalize(e, V);
            const float ndv = rf_max(rf_dot3(N, V), 0.0f);
            /* fresnel_schlick_roughness (pbr_eval.wgsl:63-66), pow(u, 5) as u^2 u^2 u */
            const float u = rf_clamp(1.0f - ndv, 0.0f, 1.0f);
            const float u2 = u * u;
            const float p5 = u2 * u2 * u;
            float Fi[3];
            for (int k = 0; k < 3; ++k) Fi[k] = F0[k] + (rf_max(1.0f - rough, F0[k]) - F0[k]) * p5;
            /* R = reflect(-V, N) = -V - 2 dot(N, -V) N */
            const float I[3] = {-V[0], -V[1], -V[2]};
            const float dni = rf_dot3(N, I);
            float R[3];
            for (int k = 0; k < 3; ++k) R[k] = I[k] - 2.0f * dni * N[k];
            float spec[3] = {0.0f, 0.0f, 0.0f};
            if (g->enable_env_reflections != 0u) {
                float env[3], brdf[2];
                or_reflection_env(g, in, P, R, rough * ENV_MAX_LOD, env);
                or_env_brdf_approx(ndv, rough, brdf);
                for (int k = 0; k < 3; ++k) spec[k] = env[k] * (F0[k] * brdf[0] + brdf[1]);
            }
            if (g->enable_reflections != 0u && ssr[3] > 0.0f)
                for (int k = 0; k < 3; ++k) spec[k] = rf_mix(spec[k], ssr[k] * Fi[k], ssr[3]);
            if (in->planar.data_d)
                rf_decode((const uint8_t*)in->planar.data_d + (size_t)y * in->planar.pitch_bytes +
                              (size_t)x * rf_texel_bytes(in->planar.format),
                          in->planar.format, pl);
            if (g->enable_reflections != 0u && pl[3] > 0.0f)
                for (int k = 0; k < 3; ++k) spec[k] = rf_mix(spec[k], pl[k] * Fi[k], pl[3]);
            if (g->has_rc_gi > 0u && rough > 0.6f) {
                /* sample_rc_specular: sample_rc_irradiance's body along
                 * normalize(mix(R, N, roughness * 0.5)) */
                float rc[3] = {0.0f, 0.0f, 0.0f};
                if (in->rc_cascade0.data_d) {
                    const float t = rough * 0.5f;
                    const float m[3] = {rf_mix(R[0], N[0], t), rf_mix(R[1], N[1], t), rf_mix(R[2], N[2], t)};
                    float L[3];
                    rf_normalize(m, L);
                    or_rc_irradiance(&in->rc_cascade0, g->rc_world_min, g->rc_world_max, P, L, rc);
                }
                const float w = rf_smoothstep(0.6f, 0.9f, rough) * 0.4f;
                for (int k = 0; k < 3; ++k) spec[k] = rf_mix(spec[k], rc[k], w);
            }
            if (flags & HL_REFLECTION_STORE) {
                for (int k = 0; k < 3; ++k) o[k] = spec[k] * ao;
                o[3] = 0.0f;
            } else {
                for (int k = 0; k < 3; ++k) o[k] = o[k] + spec[k] * ao;
            }
            ++written;
        }
    }
    return written;
}
