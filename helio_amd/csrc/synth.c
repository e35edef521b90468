/*
 * synth.c -- deterministic synthetic scenes for the deferred-lighting path:
 * camera, GpuLight arrays and G-buffers.  Host code (OpenMP), exported from
 * libhelio_synth.so.  Every random draw is a pure function of
 * (seed, stream, element, draw), so the output does not depend on threading.
 *
 * Scene model (the same for every workload, sizes from the caller):
 *   camera  glam look_at_rh + directx perspective_rh (standard Z in [0,1]),
 *           crates/helio/src/scene/camera.rs:115-125, GpuCameraUniforms::new
 *           crates/libhelio/src/camera.rs:33-58;
 *   depth   ray cast against the ground plane y = 0 and a field of boxes,
 *           view depth perturbed by N(0, depth_noise);
 *   normals uniform unit vectors flipped into the camera-facing hemisphere;
 *   albedo  uniform [0,1]; roughness U[0.05,1]; metallic 0 or 1 (p = 0.3);
 *           AO 1; emissive 0; F0 = clamp(mix(0.04, albedo, metallic), 0, 0.999)
 *           stored in normal.a / orm.a / emissive.a (gbuffer.wgsl:324-328, 488-491);
 *   lights  GpuLight with shadow_index = u32::MAX, IES/gobo = -1 and the
 *           remaining fields at GpuLight::default() (crates/libhelio/src/light.rs:117-150).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "synth.h"

#define SYN_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ RNG */
static inline uint64_t sm64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
static inline uint64_t rbits(uint64_t seed, uint32_t stream, uint64_t elem, uint32_t draw) {
    return sm64(sm64(seed ^ ((uint64_t)stream << 56)) ^ sm64(elem * 0x100000001b3ull + draw));
}
/* uniform double in [0,1) */
static inline double runif(uint64_t seed, uint32_t stream, uint64_t elem, uint32_t draw) {
    return (double)(rbits(seed, stream, elem, draw) >> 11) * (1.0 / 9007199254740992.0);
}
static inline double rnorm(uint64_t seed, uint32_t stream, uint64_t elem, uint32_t draw) {
    double u1 = runif(seed, stream, elem, draw), u2 = runif(seed, stream, elem, draw + 1);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
enum { ST_GBUF = 0, ST_LIGHTS = 1, ST_BOXES = 5 };

/* --------------------------------------------------------------- fp16 */
static uint16_t f2h(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t absx = x & 0x7fffffffu;
    if (absx >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u : 0));
    if (absx >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* overflow -> inf */
    if (absx < 0x38800000u) {                                   /* subnormal half */
        if (absx < 0x33000000u) return (uint16_t)sign;
        uint32_t e = absx >> 23, m = (absx & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126 - e; /* value / 2^-24 = m * 2^(e - 126) */
        uint32_t r = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (r & 1))) ++r;
        return (uint16_t)(sign | r);
    }
    uint32_t r = (absx - 0x38000000u) >> 13;
    uint32_t rem = absx & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1))) ++r;
    return (uint16_t)(sign | r);
}

static float h2f(uint16_t h) {
    uint32_t sign = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1f, m = h & 0x3ff, bits;
    if (e == 0) {
        float f = (float)m * (1.0f / 16777216.0f); /* exact: m * 2^-24 */
        return sign ? -f : f;
    }
    if (e == 31) bits = sign | 0x7f800000u | (m << 13);
    else bits = sign | ((e + 112) << 23) | (m << 13);
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

/* ------------------------------------------------------------- camera */
static void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static void norm3(double v[3]) {
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
}
/* 4x4 column-major, m[c*4+r] */
static void mul44(const double a[16], const double b[16], double o[16]) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
            o[c * 4 + r] = s;
        }
}
static int inv44(const double m[16], double o[16]) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = m[c * 4 + r];
            a[r][4 + c] = (r == c);
        }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
        if (fabs(a[p][c]) < 1e-300) return -1;
        for (int k = 0; k < 8; ++k) {
            double t = a[c][k];
            a[c][k] = a[p][k];
            a[p][k] = t;
        }
        double d = a[c][c];
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double f = a[r][c];
                for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
            }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) o[c * 4 + r] = a[r][4 + c];
    return 0;
}

typedef struct cam64 {
    double view[16], proj[16], vp[16], ivp[16];
    double eye[3], fwd[3], right[3], up[3];
    double tan_half, aspect;
} cam64;

static void make_cam(const hl_synth_config* c, cam64* k) {
    double eye[3] = {c->eye[0], c->eye[1], c->eye[2]};
    double f[3] = {c->target[0] - eye[0], c->target[1] - eye[1], c->target[2] - eye[2]};
    norm3(f);
    double upw[3] = {0, 1, 0};
    if (fabs(f[1]) > 0.999) { upw[1] = 0; upw[2] = -1; } /* looking straight up/down */
    double s[3];
    cross3(f, upw, s);
    norm3(s);
    double u[3];
    cross3(s, f, u);
    /* glam look_at_rh (camera.rs:115-125 uses glam::camera::rh::view::look_at_mat4) */
    double v[16] = {s[0], u[0], -f[0], 0,  s[1], u[1], -f[1], 0,  s[2], u[2], -f[2], 0,
                    -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]),
                    -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]),
                    (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]), 1};
    double fov = c->fov_y_deg * 3.141592653589793 / 180.0;
    double h = cos(0.5 * fov) / sin(0.5 * fov);
    double aspect = (double)c->width / (double)c->height;
    double w = h / aspect;
    double n = c->near_plane, fa = c->far_plane;
    double r = fa / (n - fa);
    /* glam perspective_rh (directx depth range [0,1]) */
    double p[16] = {w, 0, 0, 0, 0, h, 0, 0, 0, 0, r, -1, 0, 0, r * n, 0};
    memcpy(k->view, v, sizeof v);
    memcpy(k->proj, p, sizeof p);
    mul44(p, v, k->vp);
    inv44(k->vp, k->ivp);
    memcpy(k->eye, eye, sizeof eye);
    memcpy(k->fwd, f, sizeof f);
    memcpy(k->right, s, sizeof s);
    memcpy(k->up, u, sizeof u);
    k->tan_half = tan(0.5 * fov);
    k->aspect = aspect;
}

SYN_API int hl_synth_camera(const hl_synth_config* c, GpuCameraUniforms* out) {
    if (!c || !out || c->width == 0 || c->height == 0) return -1;
    cam64 k;
    make_cam(c, &k);
    memset(out, 0, sizeof *out);
    for (int i = 0; i < 16; ++i) {
        out->view[i] = (float)k.view[i];
        out->proj[i] = (float)k.proj[i];
        out->view_proj[i] = (float)k.vp[i];
        out->inv_view_proj[i] = (float)k.ivp[i];
        out->prev_view_proj[i] = (float)k.vp[i];
    }
    for (int i = 0; i < 3; ++i) {
        out->position_near[i] = (float)k.eye[i];
        out->forward_far[i] = (float)k.fwd[i];
    }
    out->position_near[3] = c->near_plane;
    out->forward_far[3] = c->far_plane;
    return 0;
}

/* --------------------------------------------------------------- lights */
static void default_light(GpuLight* l) {
    memset(l, 0, sizeof *l);
    l->direction_outer[1] = -1.0f;
    l->color_intensity[0] = l->color_intensity[1] = l->color_intensity[2] = l->color_intensity[3] = 1.0f;
    l->shadow_index = HL_NO_SHADOW;
    l->light_type = HL_LIGHT_POINT;
    l->god_rays_density = 1.0f;
    l->god_rays_weight = 0.6f;
    l->god_rays_decay = 1.0f;
    l->god_rays_exposure = 0.7f;
    l->flare_intensity = 1.0f;
    l->flare_scale = 1.0f;
    l->flare_tint_r = l->flare_tint_g = l->flare_tint_b = 1.0f;
    l->ies_profile_index = -1;
    l->light_function_index = -1;
    l->ies_angle_scale = 1.0f;
}

SYN_API int hl_synth_lights(const hl_synth_config* c, GpuLight* out) {
    if (!c || (!out && c->num_lights)) return -1;
    const uint32_t n = c->num_lights;
    uint32_t nspot = (uint32_t)floor((double)c->spot_fraction * n + 0.5);
    /* random slots for the spot lights: Fisher-Yates over [0, n) */
    uint8_t* is_spot = (uint8_t*)calloc(n ? n : 1, 1);
    uint32_t* perm = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    for (uint32_t i = n; i > 1; --i) {
        uint32_t j = (uint32_t)(rbits(c->seed, ST_LIGHTS, 0xFFFFFFFFull, i) % i);
        uint32_t t = perm[i - 1];
        perm[i - 1] = perm[j];
        perm[j] = t;
    }
    for (uint32_t i = 0; i < nspot && i < n; ++i) is_spot[perm[i]] = 1;
    for (uint32_t i = 0; i < n; ++i) {
        GpuLight* l = &out[i];
        default_light(l);
        for (int a = 0; a < 3; ++a)
            l->position_range[a] = (float)(c->light_box_center[a] +
                                           (runif(c->seed, ST_LIGHTS, i, a) - 0.5) * c->light_box_size[a]);
        l->position_range[3] = (float)(c->range_min + runif(c->seed, ST_LIGHTS, i, 3) * (c->range_max - c->range_min));
        for (int a = 0; a < 3; ++a)
            l->color_intensity[a] = (float)(c->color_min + runif(c->seed, ST_LIGHTS, i, 4 + a) * (c->color_max - c->color_min));
        l->color_intensity[3] =
            (float)(c->intensity_min + runif(c->seed, ST_LIGHTS, i, 7) * (c->intensity_max - c->intensity_min));
        if (is_spot[i]) {
            double z = 2.0 * runif(c->seed, ST_LIGHTS, i, 8) - 1.0;
            double phi = 6.283185307179586 * runif(c->seed, ST_LIGHTS, i, 9);
            double s = sqrt(1.0 - z * z);
            float d[3] = {(float)(s * cos(phi)), (float)(s * sin(phi)), (float)z};
            float dl = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            double outer = (c->spot_outer_min_deg +
                            runif(c->seed, ST_LIGHTS, i, 10) * (c->spot_outer_max_deg - c->spot_outer_min_deg)) *
                           3.141592653589793 / 180.0;
            double inner = outer * c->spot_inner_ratio;
            l->light_type = HL_LIGHT_SPOT;
            l->direction_outer[0] = d[0] / dl;
            l->direction_outer[1] = d[1] / dl;
            l->direction_outer[2] = d[2] / dl;
            l->direction_outer[3] = (float)cos(outer);
            l->inner_angle = (float)cos(inner);
        } else {
            l->direction_outer[0] = 0.0f;
            l->direction_outer[1] = -1.0f;
            l->direction_outer[2] = 0.0f;
            l->direction_outer[3] = 0.0f;
        }
    }
    free(is_spot);
    free(perm);
    return 0;
}

/* ------------------------------------------------------------- G-buffer */
typedef struct box {
    double lo[3], hi[3];
} box;

static double ray_box(const double o[3], const double d[3], const box* b) {
    double t0 = 0.0, t1 = 1e30;
    for (int a = 0; a < 3; ++a) {
        if (fabs(d[a]) < 1e-12) {
            if (o[a] < b->lo[a] || o[a] > b->hi[a]) return -1.0;
            continue;
        }
        double inv = 1.0 / d[a];
        double ta = (b->lo[a] - o[a]) * inv, tb = (b->hi[a] - o[a]) * inv;
        if (ta > tb) {
            double t = ta;
            ta = tb;
            tb = t;
        }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
        if (t0 > t1) return -1.0;
    }
    return t0 > 0.0 ? t0 : -1.0;
}

/* The scene's boxes on the ground plane y = 0 (caller frees). */
static box* make_boxes(const hl_synth_config* c) {
    const uint32_t nb = c->num_boxes;
    box* boxes = (box*)malloc(sizeof(box) * (nb ? nb : 1));
    for (uint32_t b = 0; b < nb; ++b) {
        double cx = (runif(c->seed, ST_BOXES, b, 0) * 2.0 - 1.0) * c->box_field_half;
        double cz = (runif(c->seed, ST_BOXES, b, 1) * 2.0 - 1.0) * c->box_field_half;
        double sx = c->box_size_min + runif(c->seed, ST_BOXES, b, 2) * (c->box_size_max - c->box_size_min);
        double sz = c->box_size_min + runif(c->seed, ST_BOXES, b, 3) * (c->box_size_max - c->box_size_min);
        double hy = c->box_height_min + runif(c->seed, ST_BOXES, b, 4) * (c->box_height_max - c->box_height_min);
        boxes[b].lo[0] = cx - 0.5 * sx;
        boxes[b].hi[0] = cx + 0.5 * sx;
        boxes[b].lo[1] = 0.0;
        boxes[b].hi[1] = hy;
        boxes[b].lo[2] = cz - 0.5 * sz;
        boxes[b].hi[2] = cz + 0.5 * sz;
    }
    return boxes;
}

/* Writes tightly packed attachments (pitch = width * texel size).
 * Formats: gbuf_format 0 -> albedo/orm RGBA8, normal/emissive RGBA16F;
 * 1 -> all RGBA16F; 2 -> all RGBA32F.  depth is f32. */
SYN_API int hl_synth_gbuffer(const hl_synth_config* c, void* albedo, void* normal, void* orm,
                             void* emissive, float* depth) {
    if (!c || !albedo || !normal || !orm || !emissive || !depth) return -1;
    cam64 k;
    make_cam(c, &k);
    GpuCameraUniforms cf;
    hl_synth_camera(c, &cf);
    const double p22 = cf.proj[10], p32 = cf.proj[14];
    const uint32_t nb = c->num_boxes;
    box* boxes = make_boxes(c);
    const uint32_t W = c->width, H = c->height, fmt = c->gbuf_format;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t yy = 0; yy < (int64_t)H; ++yy) {
        uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < W; ++x) {
            const uint64_t e = (uint64_t)y * W + x;
            double ndx = ((x + 0.5) / W) * 2.0 - 1.0, ndy = 1.0 - ((y + 0.5) / H) * 2.0;
            double dir[3];
            for (int a = 0; a < 3; ++a)
                dir[a] = k.fwd[a] + ndx * k.tan_half * k.aspect * k.right[a] + ndy * k.tan_half * k.up[a];
            norm3(dir);
            double t = -1.0;
            if (dir[1] < -1e-9) t = -k.eye[1] / dir[1];
            for (uint32_t b = 0; b < nb; ++b) {
                double tb = ray_box(k.eye, dir, &boxes[b]);
                if (tb > 0.0 && (t < 0.0 || tb < t)) t = tb;
            }
            float dval = 1.0f;
            if (t > 0.0) {
                double vd = t * (dir[0] * k.fwd[0] + dir[1] * k.fwd[1] + dir[2] * k.fwd[2]);
                vd += c->depth_noise * rnorm(c->seed, ST_GBUF, e, 0);
                if (vd < c->near_plane * 1.01) vd = c->near_plane * 1.01;
                if (vd < c->far_plane * 0.999) {
                    double dz = (p22 * (-vd) + p32) / vd;
                    dval = (float)dz;
                    if (dval >= 1.0f) dval = nextafterf(1.0f, 0.0f);
                }
            }
            depth[e] = dval;
            /* normal in the camera-facing hemisphere */
            double z = 2.0 * runif(c->seed, ST_GBUF, e, 2) - 1.0;
            double phi = 6.283185307179586 * runif(c->seed, ST_GBUF, e, 3);
            double s = sqrt(1.0 - z * z);
            double nrm[3] = {s * cos(phi), s * sin(phi), z};
            if (nrm[0] * -dir[0] + nrm[1] * -dir[1] + nrm[2] * -dir[2] < 0.0) {
                nrm[0] = -nrm[0];
                nrm[1] = -nrm[1];
                nrm[2] = -nrm[2];
            }
            double ar = runif(c->seed, ST_GBUF, e, 4), ag = runif(c->seed, ST_GBUF, e, 5),
                   ab = runif(c->seed, ST_GBUF, e, 6);
            double rough = 0.05 + 0.95 * runif(c->seed, ST_GBUF, e, 7);
            double metal = runif(c->seed, ST_GBUF, e, 8) < 0.3 ? 1.0 : 0.0;
            /* decoded values as the consumer will see them */
            float A[3], R, M;
            if (fmt == 0) {
                uint8_t* pa = (uint8_t*)albedo + 4 * e;
                uint8_t* po = (uint8_t*)orm + 4 * e;
                pa[0] = (uint8_t)floor(ar * 255.0 + 0.5);
                pa[1] = (uint8_t)floor(ag * 255.0 + 0.5);
                pa[2] = (uint8_t)floor(ab * 255.0 + 0.5);
                pa[3] = 255;
                po[0] = 255;
                po[1] = (uint8_t)floor(rough * 255.0 + 0.5);
                po[2] = metal > 0.5 ? 255 : 0;
                for (int a = 0; a < 3; ++a) A[a] = (float)pa[a] / 255.0f;
                R = (float)po[1] / 255.0f;
                M = (float)po[2] / 255.0f;
            } else {
                A[0] = (float)ar;
                A[1] = (float)ag;
                A[2] = (float)ab;
                R = (float)rough;
                M = (float)metal;
                if (fmt == 1) { /* F0 is resolved from the values the fp16 texels hold */
                    for (int a = 0; a < 3; ++a) A[a] = h2f(f2h(A[a]));
                    R = h2f(f2h(R));
                }
            }
            if (fmt == 1) {
                uint16_t* pa = (uint16_t*)albedo + 4 * e;
                uint16_t* po = (uint16_t*)orm + 4 * e;
                pa[0] = f2h(A[0]);
                pa[1] = f2h(A[1]);
                pa[2] = f2h(A[2]);
                pa[3] = f2h(1.0f);
                po[0] = f2h(1.0f);
                po[1] = f2h(R);
                po[2] = f2h(M);
            }
            float F0[3];
            for (int a = 0; a < 3; ++a) {
                float v = 0.04f * (1.0f - M) + A[a] * M;
                F0[a] = v < 0.0f ? 0.0f : (v > 0.999f ? 0.999f : v);
            }
            if (fmt == 0) {
                ((uint8_t*)orm)[4 * e + 3] = (uint8_t)floor(F0[1] * 255.0f + 0.5f);
                uint16_t* pn = (uint16_t*)normal + 4 * e;
                uint16_t* pe = (uint16_t*)emissive + 4 * e;
                for (int a = 0; a < 3; ++a) pn[a] = f2h((float)nrm[a]);
                pn[3] = f2h(F0[0]);
                pe[0] = pe[1] = pe[2] = 0;
                pe[3] = f2h(F0[2]);
            } else if (fmt == 1) {
                uint16_t* pn = (uint16_t*)normal + 4 * e;
                uint16_t* pe = (uint16_t*)emissive + 4 * e;
                ((uint16_t*)orm)[4 * e + 3] = f2h(F0[1]);
                for (int a = 0; a < 3; ++a) pn[a] = f2h((float)nrm[a]);
                pn[3] = f2h(F0[0]);
                pe[0] = pe[1] = pe[2] = 0;
                pe[3] = f2h(F0[2]);
            } else {
                float* pa = (float*)albedo + 4 * e;
                float* po = (float*)orm + 4 * e;
                float* pn = (float*)normal + 4 * e;
                float* pe = (float*)emissive + 4 * e;
                pa[0] = A[0];
                pa[1] = A[1];
                pa[2] = A[2];
                pa[3] = 1.0f;
                po[0] = 1.0f;
                po[1] = R;
                po[2] = M;
                po[3] = F0[1];
                for (int a = 0; a < 3; ++a) pn[a] = (float)nrm[a];
                pn[3] = F0[0];
                pe[0] = pe[1] = pe[2] = 0.0f;
                pe[3] = F0[2];
            }
        }
    }
    free(boxes);
    return 0;
}

/* ---------------------------------------------------------- shadow maps */
/* glam look_at_rh from eye along unit f */
static void look_along(const double eye[3], const double f_in[3], double v[16]) {
    double f[3] = {f_in[0], f_in[1], f_in[2]};
    norm3(f);
    double upw[3] = {0, 1, 0};
    if (fabs(f[1]) > 0.999) { upw[1] = 0; upw[2] = -1; }
    double s[3], u[3];
    cross3(f, upw, s);
    norm3(s);
    cross3(s, f, u);
    const double m[16] = {s[0], u[0], -f[0], 0,  s[1], u[1], -f[1], 0,  s[2], u[2], -f[2], 0,
                          -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]),
                          -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]),
                          (f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2]), 1};
    memcpy(v, m, sizeof m);
}

/* glam perspective_rh, depth [0, 1] */
static void persp(double fov, double n, double fa, double p[16]) {
    const double h = cos(0.5 * fov) / sin(0.5 * fov), r = fa / (n - fa);
    const double m[16] = {h, 0, 0, 0, 0, h, 0, 0, 0, 0, r, -1, 0, 0, r * n, 0};
    memcpy(p, m, sizeof m);
}

/* glam orthographic_rh, depth [0, 1] */
static void ortho(double hw, double n, double fa, double p[16]) {
    const double m[16] = {1.0 / hw, 0, 0, 0, 0, 1.0 / hw, 0, 0, 0, 0, 1.0 / (n - fa), 0, 0, 0, n / (n - fa), 1};
    memcpy(p, m, sizeof m);
}

/* Depth of the first surface of the scene along the texel ray of a layer,
 * as NDC z of the layer's matrix; 1 when nothing is hit. dynamic selects the
 * boxes of the dynamic atlas (odd) or the static one (even + ground). */
static float shadow_texel(const double vp[16], const double ivp[16], uint32_t size, uint32_t x, uint32_t y,
                          const box* boxes, uint32_t nb, int dynamic) {
    const double ndx = ((x + 0.5) / size) * 2.0 - 1.0, ndy = 1.0 - ((y + 0.5) / size) * 2.0;
    double a[3], b[3];
    for (int k = 0; k < 2; ++k) {
        const double z = k == 0 ? 0.0 : 1.0;
        double h[4];
        for (int r = 0; r < 4; ++r) h[r] = ivp[r] * ndx + ivp[4 + r] * ndy + ivp[8 + r] * z + ivp[12 + r];
        double* o = k == 0 ? a : b;
        for (int r = 0; r < 3; ++r) o[r] = h[r] / h[3];
    }
    double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int r = 0; r < 3; ++r) d[r] /= len;
    double t = -1.0;
    if (!dynamic && fabs(d[1]) > 1e-12) {
        const double tg = -a[1] / d[1];
        if (tg > 0.0) t = tg;
    }
    for (uint32_t i = (uint32_t)dynamic; i < nb; i += 2) {
        const double tb = ray_box(a, d, &boxes[i]);
        if (tb > 0.0 && (t < 0.0 || tb < t)) t = tb;
    }
    if (t < 0.0 || t > len) return 1.0f;
    const double p[4] = {a[0] + t * d[0], a[1] + t * d[1], a[2] + t * d[2], 1.0};
    double cz = 0.0, cw = 0.0;
    for (int c = 0; c < 4; ++c) {
        cz += vp[c * 4 + 2] * p[c];
        cw += vp[c * 4 + 3] * p[c];
    }
    const double z = cz / cw;
    return z < 0.0 ? 0.0f : (z >= 1.0 ? 1.0f : (float)z);
}

/* Shadow atlases (dynamic: odd boxes; static: even boxes and the ground) and
 * one light matrix per layer for the lights with a shadow_index: point lights
 * 6 cube faces (+X, -X, +Y, -Y, +Z, -Z; 90 degrees, near 0.05, far = range),
 * spot lights one perspective layer (the outer cone plus 5 degrees), the
 * directional light 4 orthographic cascades centred on the camera eye with
 * half-widths csm[c].  Layers are size x size, layer-major. */
SYN_API int hl_synth_shadow_maps(const hl_synth_config* c, const GpuLight* lights, uint32_t num_lights,
                                 const float csm[4], uint32_t size, uint32_t num_layers, float* atlas,
                                 float* static_atlas, float* matrices) {
    if (!c || !atlas || !static_atlas || !matrices || size == 0) return -1;
    box* boxes = make_boxes(c);
    const uint32_t nb = c->num_boxes;
    const size_t texels = (size_t)size * size;
    for (size_t i = 0; i < texels * num_layers; ++i) atlas[i] = static_atlas[i] = 1.0f;
    memset(matrices, 0, sizeof(float) * 16 * num_layers);
    static const double faces[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (uint32_t li = 0; li < num_lights; ++li) {
        const GpuLight* l = &lights[li];
        if (l->shadow_index == HL_NO_SHADOW) continue;
        const double pos[3] = {l->position_range[0], l->position_range[1], l->position_range[2]};
        const double dir[3] = {l->direction_outer[0], l->direction_outer[1], l->direction_outer[2]};
        uint32_t nl = l->light_type == HL_LIGHT_POINT ? 6u : (l->light_type == HL_LIGHT_DIRECTIONAL ? 4u : 1u);
        for (uint32_t k = 0; k < nl; ++k) {
            const uint32_t layer = l->shadow_index + k;
            if (layer >= num_layers) break;
            double v[16], p[16], vp[16], ivp[16];
            if (l->light_type == HL_LIGHT_POINT) {
                look_along(pos, faces[k], v);
                persp(3.141592653589793 / 2.0, 0.05, l->position_range[3], p);
            } else if (l->light_type == HL_LIGHT_DIRECTIONAL) {
                const double hw = csm[k];
                const double eye[3] = {c->eye[0] - dir[0] * 500.0, c->eye[1] - dir[1] * 500.0,
                                       c->eye[2] - dir[2] * 500.0};
                look_along(eye, dir, v);
                ortho(hw, 0.0, 1000.0, p);
            } else {
                double fov = 2.0 * acos(l->direction_outer[3]) + 10.0 * 3.141592653589793 / 180.0;
                if (fov > 3.0) fov = 3.0;
                look_along(pos, dir, v);
                persp(fov, 0.05, l->position_range[3], p);
            }
            mul44(p, v, vp);
            inv44(vp, ivp);
            for (int e = 0; e < 16; ++e) matrices[16 * (size_t)layer + e] = (float)vp[e];
            float* da = atlas + texels * layer;
            float* sa = static_atlas + texels * layer;
#pragma omp parallel for schedule(dynamic, 8)
            for (int64_t yy = 0; yy < (int64_t)size; ++yy)
                for (uint32_t x = 0; x < size; ++x) {
                    da[(size_t)yy * size + x] = shadow_texel(vp, ivp, size, x, (uint32_t)yy, boxes, nb, 1);
                    sa[(size_t)yy * size + x] = shadow_texel(vp, ivp, size, x, (uint32_t)yy, boxes, nb, 0);
                }
        }
    }
    free(boxes);
    return 0;
}

/* Uniform random subset of [0, pool) of size n, ascending (C5 view light sets). */
SYN_API int hl_synth_subset(uint64_t seed, uint32_t view, uint32_t pool, uint32_t n, uint32_t* out) {
    if (n > pool || (!out && n)) return -1;
    /* selection sampling (Knuth, Algorithm S): ascending by construction */
    uint32_t chosen = 0;
    for (uint32_t i = 0; i < pool && chosen < n; ++i) {
        double u = runif(seed, 4, ((uint64_t)view << 32) | i, 0);
        if ((double)(pool - i) * u < (double)(n - chosen)) out[chosen++] = i;
    }
    return 0;
}

SYN_API uint32_t hl_synth_abi_size(void) { return (uint32_t)sizeof(hl_synth_config); }
