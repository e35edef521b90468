/*
 * oracle.c -- FP32 restatement of the deferred-lighting hot path (test
 * infrastructure only; see oracle.h).  Compiled with -ffp-contract=off so every
 * line below is one IEEE-754 binary32 operation, rounded to nearest.
 */
/* gcc: -ffp-contract=off (set by the build) keeps every operation separately rounded */

#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------ small math */

/* pbr_eval.wgsl:6 -- PI: f32 = 3.14159265359 */
static const float OR_PI = 3.14159265359f;

static inline float or_dot(const float a[3], const float b[3]) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}
static inline float or_length(const float v[3]) { return sqrtf(or_dot(v, v)); }
static inline void or_normalize(float v[3]) {
    float l = or_length(v);
    v[0] = v[0] / l;
    v[1] = v[1] / l;
    v[2] = v[2] / l;
}
static inline float or_maxf(float a, float b) { return a > b ? a : b; }  /* finite inputs */
static inline float or_minf(float a, float b) { return a < b ? a : b; }
static inline float or_clampf(float x, float lo, float hi) { return or_minf(or_maxf(x, lo), hi); }

/* mat4 (column-major) * vec4, columns accumulated in order. */
static inline void or_mat4_mul(const float m[16], const float v[4], float r[4]) {
    for (int row = 0; row < 4; ++row) {
        float s = m[0 * 4 + row] * v[0];
        s = s + m[1 * 4 + row] * v[1];
        s = s + m[2 * 4 + row] * v[2];
        s = s + m[3 * 4 + row] * v[3];
        r[row] = s;
    }
}

/* WGSL smoothstep(e0, e1, x). */
static inline float or_smoothstep(float e0, float e1, float x) {
    float t = or_clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

/* pbr_eval.wgsl:8 */
static inline float or_pow5(float x) {
    float x2 = x * x;
    return x2 * x2 * x;
}

/* ---------------------------------------------------------- texel decode */

static float or_half_to_float(uint16_t h) {
    uint32_t sign = (uint32_t)(h >> 15) << 31;
    uint32_t e = (h >> 10) & 0x1f;
    uint32_t m = h & 0x3ff;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = sign;
        } else { /* subnormal half -> normal float */
            int shift = 0;
            while ((m & 0x400) == 0) { m <<= 1; ++shift; }
            m &= 0x3ff;
            bits = sign | ((uint32_t)(127 - 15 + 1 - shift) << 23) | (m << 13);
        }
    } else if (e == 31) {
        bits = sign | 0x7f800000u | (m << 13);
    } else {
        bits = sign | ((e + 127 - 15) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

/* textureLoad(tex, (x, y), 0) as vec4<f32>. */
static void or_load_rgba(const hl_texture* t, uint32_t x, uint32_t y, float out[4]) {
    const uint8_t* row = (const uint8_t*)t->data_d + (size_t)y * t->pitch_bytes;
    switch (t->format) {
        case HL_FMT_RGBA8_UNORM: {
            const uint8_t* p = row + (size_t)x * 4;
            for (int k = 0; k < 4; ++k) out[k] = (float)p[k] / 255.0f;
            break;
        }
        case HL_FMT_RGBA16_FLOAT: {
            const uint16_t* p = (const uint16_t*)(row + (size_t)x * 8);
            for (int k = 0; k < 4; ++k) out[k] = or_half_to_float(p[k]);
            break;
        }
        case HL_FMT_RGBA32_FLOAT: {
            const float* p = (const float*)(row + (size_t)x * 16);
            for (int k = 0; k < 4; ++k) out[k] = p[k];
            break;
        }
        default:
            out[0] = out[1] = out[2] = out[3] = 0.0f;
    }
}

static float or_load_depth(const hl_texture* t, uint32_t x, uint32_t y) {
    const uint8_t* row = (const uint8_t*)t->data_d + (size_t)y * t->pitch_bytes;
    float d;
    memcpy(&d, row + (size_t)x * 4, 4);
    return d;
}

/* ------------------------------------------------------------ tile cull */

/* light_cull.wgsl:85-100 make_plane_from_ndc_edge */
static void or_plane_from_ndc_edge(const GpuCameraUniforms* cam, float p0x, float p0y, float p1x,
                                   float p1y, float plane[4]) {
    const float p00 = cam->proj[0];  /* proj[0][0] */
    const float p11 = cam->proj[5];  /* proj[1][1] */
    float r0[3] = {p0x / p00, p0y / p11, -1.0f};
    float r1[3] = {p1x / p00, p1y / p11, -1.0f};
    or_normalize(r0);
    or_normalize(r1);
    float n[3] = {r0[1] * r1[2] - r0[2] * r1[1], r0[2] * r1[0] - r0[0] * r1[2],
                  r0[0] * r1[1] - r0[1] * r1[0]};
    or_normalize(n);
    plane[0] = n[0];
    plane[1] = n[1];
    plane[2] = n[2];
    plane[3] = 0.0f;
}

/* light_cull.wgsl:105-117 sphere_inside_tile_frustum */
static int or_sphere_inside(const float c[3], float r, const float planes[4][4]) {
    for (int i = 0; i < 4; ++i) {
        float d = or_dot(planes[i], c) + planes[i][3];
        if (d < -r) return 0;
    }
    return 1;
}

/* light_cull.wgsl:136-151: the tile's NDC rectangle and its four planes
 * (left, right, top, bottom). */
static void or_tile_planes(const GpuCameraUniforms* cam, uint32_t tile_x, uint32_t tile_y,
                           uint32_t width, uint32_t height, float planes[4][4]) {
    const float fw = (float)width, fh = (float)height, ts = (float)HL_TILE_SIZE;
    float ndc_left = ((float)tile_x * ts / fw) * 2.0f - 1.0f;
    float ndc_right = or_minf(((float)(tile_x + 1u) * ts / fw) * 2.0f - 1.0f, 1.0f);
    float ndc_top = 1.0f - ((float)tile_y * ts / fh) * 2.0f;
    float ndc_bottom = or_maxf(1.0f - ((float)(tile_y + 1u) * ts / fh) * 2.0f, -1.0f);
    or_plane_from_ndc_edge(cam, ndc_left, ndc_bottom, ndc_left, ndc_top, planes[0]);
    or_plane_from_ndc_edge(cam, ndc_right, ndc_top, ndc_right, ndc_bottom, planes[1]);
    or_plane_from_ndc_edge(cam, ndc_left, ndc_top, ndc_right, ndc_top, planes[2]);
    or_plane_from_ndc_edge(cam, ndc_right, ndc_bottom, ndc_left, ndc_bottom, planes[3]);
}

/* (view * vec4(pos, 1)).xyz -- light_cull.wgsl:168 */
static void or_light_view_pos(const GpuCameraUniforms* cam, const GpuLight* l, float out[3]) {
    float v[4] = {l->position_range[0], l->position_range[1], l->position_range[2], 1.0f};
    float r[4];
    or_mat4_mul(cam->view, v, r);
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
}

void or_light_cull_tiles(const GpuCameraUniforms* cam, const LightCullParams* params,
                         const GpuLight* lights, uint32_t* lists, uint32_t* counts) {
    const uint32_t total = params->num_tiles_x * params->num_tiles_y;
    if (params->num_lights == 0) { /* LightCullPass::execute clears both (lib.rs:285-292) */
        memset(lists, 0, (size_t)total * HL_MAX_LIGHTS_PER_TILE * 4);
        memset(counts, 0, (size_t)total * 4);
        return;
    }
    /* One view-space position per light; the shader recomputes it per tile
     * with the same operations, so hoisting it changes no bit. */
    float* vpos = (float*)malloc(sizeof(float) * 3 * (size_t)params->num_lights);
    for (uint32_t i = 0; i < params->num_lights; ++i) or_light_view_pos(cam, &lights[i], vpos + 3 * i);

#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < (int64_t)total; ++t) {
        uint32_t tile_idx = (uint32_t)t;
        uint32_t tile_x = tile_idx % params->num_tiles_x;
        uint32_t tile_y = tile_idx / params->num_tiles_x;
        float planes[4][4];
        or_tile_planes(cam, tile_x, tile_y, params->screen_width, params->screen_height, planes);
        uint32_t count = 0;
        for (uint32_t i = 0; i < params->num_lights; ++i) {
            const GpuLight* l = &lights[i];
            if (l->light_type == 0u) {
                if (count < HL_MAX_LIGHTS_PER_TILE) lists[(size_t)tile_idx * HL_MAX_LIGHTS_PER_TILE + count++] = i;
                continue;
            }
            if (or_sphere_inside(vpos + 3 * i, l->position_range[3], planes)) {
                if (count < HL_MAX_LIGHTS_PER_TILE) lists[(size_t)tile_idx * HL_MAX_LIGHTS_PER_TILE + count++] = i;
            }
        }
        counts[tile_idx] = count;
    }
    free(vpos);
}

/* ---------------------------------------------------------- cluster cull */

uint32_t or_slice_of_depth(const hl_cluster_params* cp, float depth) {
    uint32_t k = 0;
    for (uint32_t s = 1; s < cp->grid_z; ++s)
        if (cp->depth_split[s] <= depth) k = s;
    return k;
}

/* Lateral planes of a cluster column: the column's pixel rectangle in NDC,
 * then the same plane construction as the reference tile cull. */
static int or_column_planes(const GpuCameraUniforms* cam, const hl_cluster_params* cp, uint32_t cx,
                            uint32_t cy, float planes[4][4], float ndc[4]) {
    uint32_t px0 = cx * cp->cell_w, py0 = cy * cp->cell_h;
    if (px0 >= cp->width || py0 >= cp->height) return 0;
    uint32_t px1 = px0 + cp->cell_w, py1 = py0 + cp->cell_h;
    if (px1 > cp->width) px1 = cp->width;
    if (py1 > cp->height) py1 = cp->height;
    const float fw = (float)cp->width, fh = (float)cp->height;
    float l = ((float)px0 / fw) * 2.0f - 1.0f;
    float r = ((float)px1 / fw) * 2.0f - 1.0f;
    float t = 1.0f - ((float)py0 / fh) * 2.0f;
    float b = 1.0f - ((float)py1 / fh) * 2.0f;
    or_plane_from_ndc_edge(cam, l, b, l, t, planes[0]);
    or_plane_from_ndc_edge(cam, r, t, r, b, planes[1]);
    or_plane_from_ndc_edge(cam, l, t, r, t, planes[2]);
    or_plane_from_ndc_edge(cam, r, b, l, b, planes[3]);
    ndc[0] = l;
    ndc[1] = r;
    ndc[2] = b;
    ndc[3] = t;
    return 1;
}

/* View-space AABB of slice k of a column: the four NDC corners unprojected at
 * the padded near and far view distances. */
static void or_cluster_aabb(const GpuCameraUniforms* cam, const hl_cluster_params* cp,
                            const float ndc[4], uint32_t k, float lo[3], float hi[3]) {
    const float p00 = cam->proj[0], p11 = cam->proj[5];
    const float zs[2] = {cp->slice_near[k], cp->slice_far[k]};
    lo[0] = lo[1] = INFINITY;
    hi[0] = hi[1] = -INFINITY;
    for (int zi = 0; zi < 2; ++zi) {
        for (int xi = 0; xi < 2; ++xi) {
            float x = ndc[xi] * zs[zi] / p00;
            lo[0] = or_minf(lo[0], x);
            hi[0] = or_maxf(hi[0], x);
        }
        for (int yi = 0; yi < 2; ++yi) {
            float y = ndc[2 + yi] * zs[zi] / p11;
            lo[1] = or_minf(lo[1], y);
            hi[1] = or_maxf(hi[1], y);
        }
    }
    lo[2] = -zs[1];
    hi[2] = -zs[0];
}

static int or_sphere_aabb(const float c[3], float r, const float lo[3], const float hi[3]) {
    float d2 = 0.0f;
    for (int a = 0; a < 3; ++a) {
        float d = 0.0f;
        if (c[a] < lo[a]) d = lo[a] - c[a];
        else if (c[a] > hi[a]) d = c[a] - hi[a];
        d2 = d2 + d * d;
    }
    return d2 <= r * r;
}

/* Cone (apex c, axis a (view space, not normalized), cos = outer) against the
 * bounding sphere of the AABB, radius grown by pad.  1 = may intersect. */
static int or_cone_aabb(const float c[3], const float a[3], float outer, float pad,
                        const float lo[3], const float hi[3]) {
    float alen = or_length(a);
    if (!(alen > 0.0f)) return 1;
    float cos_t = outer / alen;
    if (!(cos_t > 0.0f)) return 1; /* half-angle >= 90 degrees: no cone test */
    if (cos_t > 1.0f) cos_t = 1.0f;
    float sin_t = sqrtf(or_maxf(0.0f, 1.0f - cos_t * cos_t));
    float m[3], ext[3];
    for (int i = 0; i < 3; ++i) {
        m[i] = (lo[i] + hi[i]) * 0.5f;
        ext[i] = hi[i] - lo[i];
    }
    float R = or_length(ext) * 0.5f + pad;
    float v[3] = {m[0] - c[0], m[1] - c[1], m[2] - c[2]};
    float an[3] = {a[0] / alen, a[1] / alen, a[2] / alen};
    float vlen2 = or_dot(v, v);
    float v1 = or_dot(v, an);
    float perp = sqrtf(or_maxf(0.0f, vlen2 - v1 * v1));
    float dcp = cos_t * perp - v1 * sin_t;
    if (dcp > R) return 0;
    if (v1 < -R) return 0;
    return 1;
}

/* View-space axis of a spot light: the rotation part of view applied to direction. */
static void or_light_view_dir(const GpuCameraUniforms* cam, const GpuLight* l, float out[3]) {
    float v[4] = {l->direction_outer[0], l->direction_outer[1], l->direction_outer[2], 0.0f};
    float r[4];
    or_mat4_mul(cam->view, v, r);
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
}

/* Membership of light l in slice k of a column (planes/ndc precomputed). */
static int or_cluster_member(const GpuCameraUniforms* cam, const hl_cluster_params* cp,
                             const GpuLight* l, const float vpos[3], const float planes[4][4],
                             const float ndc[4], uint32_t k) {
    if (l->light_type == 0u) return 1;
    const float r = l->position_range[3] + cp->slice_r_pad[k];
    if (!or_sphere_inside(vpos, r, planes)) return 0;
    float lo[3], hi[3];
    or_cluster_aabb(cam, cp, ndc, k, lo, hi);
    if (!or_sphere_aabb(vpos, r, lo, hi)) return 0;
    if (l->light_type == 2u) {
        float a[3];
        or_light_view_dir(cam, l, a);
        if (!or_cone_aabb(vpos, a, l->direction_outer[3], cp->slice_r_pad[k], lo, hi)) return 0;
    }
    return 1;
}

uint32_t or_cluster_hits(const GpuCameraUniforms* cam, const hl_cluster_params* cp,
                         const GpuLight* lights, uint32_t num_lights, uint32_t cluster) {
    uint32_t cx = cluster % cp->grid_x;
    uint32_t cy = (cluster / cp->grid_x) % cp->grid_y;
    uint32_t cz = cluster / (cp->grid_x * cp->grid_y);
    float planes[4][4], ndc[4];
    if (!or_column_planes(cam, cp, cx, cy, planes, ndc)) return 0;
    uint32_t hits = 0;
    for (uint32_t i = 0; i < num_lights; ++i) {
        float vpos[3];
        or_light_view_pos(cam, &lights[i], vpos);
        hits += (uint32_t)or_cluster_member(cam, cp, &lights[i], vpos, planes, ndc, cz);
    }
    return hits;
}

uint32_t or_light_cull_clusters(const GpuCameraUniforms* cam, const hl_cluster_params* cp,
                                const GpuLight* lights, uint32_t num_lights, uint32_t* offsets,
                                uint32_t* counts, uint32_t* pool, uint32_t pool_cap) {
    const uint32_t ncol = cp->grid_x * cp->grid_y;
    const uint32_t ncell = ncol * cp->grid_z;
    const uint32_t cap = cp->cap;
    float* vpos = (float*)malloc(sizeof(float) * 3 * (size_t)(num_lights ? num_lights : 1));
    for (uint32_t i = 0; i < num_lights; ++i) or_light_view_pos(cam, &lights[i], vpos + 3 * i);

    /* Pass 1: capped counts (lists kept in a per-cluster scratch). */
    uint32_t** tmp = (uint32_t**)calloc(ncell, sizeof(uint32_t*));
    uint32_t* raw = (uint32_t*)calloc(ncell, sizeof(uint32_t));
    uint32_t flags = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : flags)
    for (int64_t col = 0; col < (int64_t)ncol; ++col) {
        uint32_t cx = (uint32_t)col % cp->grid_x, cy = (uint32_t)col / cp->grid_x;
        float planes[4][4], ndc[4];
        int valid = or_column_planes(cam, cp, cx, cy, planes, ndc);
        for (uint32_t k = 0; k < cp->grid_z; ++k) {
            uint32_t c = (uint32_t)col + ncol * k;
            uint32_t n = 0;
            uint32_t* list = NULL;
            if (valid) {
                list = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(num_lights ? num_lights : 1));
                for (uint32_t i = 0; i < num_lights; ++i)
                    if (or_cluster_member(cam, cp, &lights[i], vpos + 3 * i, planes, ndc, k)) {
                        if (cap == 0 || n < cap) list[n] = i;
                        ++n;
                    }
            }
            if (cap != 0 && n > cap) {
                flags |= HL_OVERFLOW_CAP;
                n = cap;
            }
            tmp[c] = list;
            raw[c] = n;
        }
    }
    /* Exclusive scan, pool truncation, fill. */
    uint64_t off = 0;
    for (uint32_t c = 0; c < ncell; ++c) {
        uint32_t o = off > pool_cap ? pool_cap : (uint32_t)off;
        uint32_t n = raw[c];
        uint32_t room = pool_cap - o;
        if (n > room) {
            flags |= HL_OVERFLOW_POOL;
            n = room;
        }
        offsets[c] = o;
        counts[c] = n;
        if (n) memcpy(pool + o, tmp[c], sizeof(uint32_t) * n);
        off += raw[c];
        free(tmp[c]);
    }
    free(tmp);
    free(raw);
    free(vpos);
    return flags;
}

/* ------------------------------------------------------------- shading */

/* deferred_lighting.wgsl:699-820 pbr_direct_light with sf = 1 (no shadow
 * caster bound), no IES/gobo (indices < 0), isotropic, no subsurface. */
/* pbr_direct_light with its shadow factor sf (1 for the unshadowed call). */
static void or_pbr_direct_light_sf(const GpuLight* light, const float wp[3], const float N[3],
                                   const float V[3], const float F0[3], const float albedo[3],
                                   float roughness, float metallic, float sf, float out[3]) {
    float L[3], radiance[3];
    out[0] = out[1] = out[2] = 0.0f;
    if (light->light_type == 0u) {
        L[0] = -light->direction_outer[0];
        L[1] = -light->direction_outer[1];
        L[2] = -light->direction_outer[2];
        or_normalize(L);
        for (int k = 0; k < 3; ++k) radiance[k] = light->color_intensity[k] * light->color_intensity[3];
    } else {
        float to_light[3] = {light->position_range[0] - wp[0], light->position_range[1] - wp[1],
                             light->position_range[2] - wp[2]};
        float dist = or_length(to_light);
        if (dist > light->position_range[3]) return;
        L[0] = to_light[0] / dist;
        L[1] = to_light[1] / dist;
        L[2] = to_light[2] / dist;
        float atten = 1.0f / (dist * dist + 0.0001f);
        float nd = dist / light->position_range[3];
        atten = atten * or_maxf(0.0f, 1.0f - nd * nd * nd * nd);
        if (light->light_type == 2u) {
            float negL[3] = {-L[0], -L[1], -L[2]};
            float cos_a = or_dot(negL, light->direction_outer);
            atten = atten * or_smoothstep(light->direction_outer[3], light->inner_angle, cos_a);
        }
        for (int k = 0; k < 3; ++k)
            radiance[k] = light->color_intensity[k] * light->color_intensity[3] * atten;
    }
    float NdL = or_maxf(or_dot(N, L), 0.0f);
    if (NdL == 0.0f) return;
    if (radiance[0] < 0.002f && radiance[1] < 0.002f && radiance[2] < 0.002f) return;

    float H[3] = {V[0] + L[0], V[1] + L[1], V[2] + L[2]};
    or_normalize(H);
    float HdV = or_maxf(or_dot(H, V), 0.0f);
    float p5 = or_pow5(or_clampf(1.0f - HdV, 0.0f, 1.0f));
    float F[3], kD[3];
    for (int k = 0; k < 3; ++k) {
        F[k] = F0[k] + (1.0f - F0[k]) * p5;
        kD[k] = (1.0f - F[k]) * (1.0f - metallic);
    }
    float NdV = or_maxf(or_dot(N, V), 0.0f);
    /* distribution_ggx, pbr_eval.wgsl:12-18 */
    float a = roughness * roughness;
    float a2 = a * a;
    float NdH = or_maxf(or_dot(N, H), 0.0f);
    float denom = NdH * NdH * (a2 - 1.0f) + 1.0f;
    float D = a2 / (OR_PI * denom * denom + 0.0001f);
    /* geometry_smith, pbr_eval.wgsl:33-43 */
    float rr = roughness + 1.0f;
    float kk = (rr * rr) / 8.0f;
    float gv = NdV / (NdV * (1.0f - kk) + kk + 0.0001f);
    float gl = NdL / (NdL * (1.0f - kk) + kk + 0.0001f);
    float G = gv * gl;
    float sden = 4.0f * NdV * NdL + 0.0001f;
    for (int k = 0; k < 3; ++k) {
        float spec = D * G * F[k] / sden;
        float diff = kD[k] * albedo[k] / OR_PI;
        out[k] = (diff + spec) * radiance[k] * NdL * sf;
    }
}

void or_pbr_direct_light(const GpuLight* light, const float wp[3], const float N[3],
                         const float V[3], const float F0[3], const float albedo[3],
                         float roughness, float metallic, float out[3]) {
    or_pbr_direct_light_sf(light, wp, N, V, F0, albedo, roughness, metallic, 1.0f, out);
}

/* Pixel -> cell list. */
static void or_cell_list(const hl_light_grid* grid, const DeferredGlobals* g, uint32_t x, uint32_t y,
                         float depth, const uint32_t** list, uint32_t* count) {
    if (grid->mode == HL_GRID_TILES) {
        /* deferred_lighting.wgsl:1121-1125 */
        uint32_t tile_x = x / HL_TILE_SIZE, tile_y = y / HL_TILE_SIZE;
        uint32_t tile_idx = tile_y * g->num_tiles_x + tile_x;
        *count = grid->counts_d[tile_idx];
        *list = grid->lists_d + (size_t)tile_idx * HL_MAX_LIGHTS_PER_TILE;
    } else {
        const hl_cluster_params* cp = grid->clusters;
        uint32_t cx = x / cp->cell_w, cy = y / cp->cell_h;
        uint32_t cz = or_slice_of_depth(cp, depth);
        uint32_t c = cx + cp->grid_x * (cy + cp->grid_y * cz);
        *count = grid->counts_d[c];
        *list = grid->lists_d + grid->offsets_d[c];
    }
}

/* Everything fs_main computes before the light loop. */
typedef struct or_px {
    float wp[3], N[3], V[3], F0[3], albedo[3], emissive[3];
    float alpha, rough, metal, ao;
} or_px;

static void or_setup_pixel(const GpuCameraUniforms* cam, const hl_gbuffer* gb, uint32_t x, uint32_t y,
                           float depth, or_px* p) {
    float albedo_a[4], normal_r[4], orm_r[4], emissive_r[4];
    or_load_rgba(&gb->albedo, x, y, albedo_a);
    or_load_rgba(&gb->normal, x, y, normal_r);
    or_load_rgba(&gb->orm, x, y, orm_r);
    or_load_rgba(&gb->emissive, x, y, emissive_r);
    for (int k = 0; k < 3; ++k) {
        p->albedo[k] = albedo_a[k];
        p->emissive[k] = emissive_r[k];
        p->N[k] = normal_r[k];
    }
    p->alpha = albedo_a[3];
    or_normalize(p->N);
    p->ao = orm_r[0] * 1.0f; /* ao * ssao (1x1 white fallback) */
    p->rough = orm_r[1];
    p->metal = orm_r[2];
    /* deferred_lighting.wgsl:1043-1047 */
    const float clip_x = (float)x + 0.5f, clip_y = (float)y + 0.5f;
    const float u = clip_x / (float)gb->width, v = clip_y / (float)gb->height;
    float ndc4[4] = {u * 2.0f - 1.0f, 1.0f - v * 2.0f, depth, 1.0f};
    float wh[4];
    or_mat4_mul(cam->inv_view_proj, ndc4, wh);
    p->wp[0] = wh[0] / wh[3];
    p->wp[1] = wh[1] / wh[3];
    p->wp[2] = wh[2] / wh[3];
    /* :1084-1086 */
    p->F0[0] = or_clampf(normal_r[3], 0.0f, 0.999f);
    p->F0[1] = or_clampf(orm_r[3], 0.0f, 0.999f);
    p->F0[2] = or_clampf(emissive_r[3], 0.0f, 0.999f);
    for (int k = 0; k < 3; ++k) p->V[k] = cam->position_near[k] - p->wp[k];
    or_normalize(p->V);
}

void or_world_pos(const GpuCameraUniforms* cam, uint32_t width, uint32_t height, uint32_t x,
                  uint32_t y, float depth, float wp[3]) {
    const float u = ((float)x + 0.5f) / (float)width, v = ((float)y + 0.5f) / (float)height;
    float ndc4[4] = {u * 2.0f - 1.0f, 1.0f - v * 2.0f, depth, 1.0f};
    float wh[4];
    or_mat4_mul(cam->inv_view_proj, ndc4, wh);
    wp[0] = wh[0] / wh[3];
    wp[1] = wh[1] / wh[3];
    wp[2] = wh[2] / wh[3];
}

static int or_debug_supported(uint32_t m) {
    return m == 0 || m == 1 || m == 2 || m == 4 || m == 5 || m == 20 || m == 21;
}

uint64_t or_shade_ex(const GpuCameraUniforms* cam, const DeferredGlobals* g, const hl_gbuffer* gb,
                     const GpuLight* lights, uint32_t num_lights, const hl_light_grid* grid,
                     const hl_shadow_maps* sh, const hl_texture* rc, float* out, uint32_t out_pitch,
                     uint32_t flags, uint32_t row0, uint32_t row1) {
    (void)num_lights;
    const int use_rc = rc && rc->data_d && g->has_rc_gi != 0u;
    if (!or_debug_supported(g->debug_mode)) return 0;
    if (row1 == 0 || row1 > gb->height) row1 = gb->height;
    uint64_t pairs = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : pairs)
    for (int64_t yy = (int64_t)row0; yy < (int64_t)row1; ++yy) {
        const uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < gb->width; ++x) {
            float* o = out ? (float*)((uint8_t*)out + (size_t)y * out_pitch) + 4 * (size_t)x : NULL;
            const float depth = or_load_depth(&gb->depth, x, y);
            if (depth >= 1.0f) { /* :996-997 discard */
                if (o && (flags & HL_SHADE_CLEAR_SKY)) o[0] = o[1] = o[2] = o[3] = 0.0f;
                continue;
            }
            or_px p;
            or_setup_pixel(cam, gb, x, y, depth, &p);
            const uint32_t dm = g->debug_mode;
            if (dm == 1 || dm == 2 || dm == 4 || dm == 20 || dm == 21) { /* :1032-1035 */
                if (o) {
                    o[0] = p.albedo[0];
                    o[1] = p.albedo[1];
                    o[2] = p.albedo[2];
                    o[3] = p.alpha;
                }
                continue;
            }
            if (dm == 5) { /* :1036-1038 */
                if (o) {
                    for (int k = 0; k < 3; ++k) o[k] = p.N[k] * 0.5f + 0.5f;
                    o[3] = 1.0f;
                }
                continue;
            }
            /* :1119-1146 direct lighting over the cell list */
            float Lo[3] = {0.0f, 0.0f, 0.0f};
            const uint32_t* list;
            uint32_t cnt;
            or_cell_list(grid, g, x, y, depth, &list, &cnt);
            for (uint32_t i = 0; i < cnt; ++i) {
                const GpuLight* l = &lights[list[i]];
                if (l->light_type != 0u) {
                    float tl[3] = {l->position_range[0] - p.wp[0], l->position_range[1] - p.wp[1],
                                   l->position_range[2] - p.wp[2]};
                    if (or_length(tl) > l->position_range[3]) continue;
                }
                ++pairs;
                if (!o) continue;
                /* :1136-1143 shadow_factor for every light in range */
                const float sf = sh ? or_shadow_factor(sh, lights, list[i], cam->position_near, g->csm_splits, p.wp,
                                                       p.N, (float)x + 0.5f, (float)y + 0.5f)
                                    : 1.0f;
                float c[3];
                or_pbr_direct_light_sf(l, p.wp, p.N, p.V, p.F0, p.albedo, p.rough, p.metal, sf, c);
                Lo[0] = Lo[0] + c[0];
                Lo[1] = Lo[1] + c[1];
                Lo[2] = Lo[2] + c[2];
            }
            if (!o) continue;
            /* :1159-1163 RC indirect diffuse */
            float rc_irr[3] = {0.0f, 0.0f, 0.0f}, diff_ind[3] = {0.0f, 0.0f, 0.0f}, rc_w = 0.0f;
            if (use_rc) {
                or_rc_irradiance(rc, g->rc_world_min, g->rc_world_max, p.wp, p.N, rc_irr);
                const float ndv = or_maxf(or_dot(p.N, p.V), 0.0f);
                const float p5 = or_pow5(or_clampf(1.0f - ndv, 0.0f, 1.0f));
                for (int k = 0; k < 3; ++k) {
                    /* fresnel_schlick_roughness, pbr_eval.wgsl:63-66 */
                    const float f = p.F0[k] + (or_maxf(1.0f - p.rough, p.F0[k]) - p.F0[k]) * p5;
                    const float kd = (1.0f - f) * (1.0f - p.metal);
                    diff_ind[k] = kd * rc_irr[k] * p.albedo[k];
                }
                /* :1198 clamp(length(rc_irr) * 4, 0, 1) */
                rc_w = or_clampf(or_length(rc_irr) * 4.0f, 0.0f, 1.0f);
            }
            /* :1189-1231 ambient + combine (no lightmap, no SSS) */
            float hemi_t = p.N[1] * 0.5f + 0.5f;
            for (int k = 0; k < 3; ++k) {
                float sky = g->ambient_color[k] * g->ambient_intensity;
                float ground = sky * 0.15f;
                float hemi = (ground * (1.0f - hemi_t) + sky * hemi_t) * p.albedo[k];
                if (use_rc) hemi = hemi * (1.0f - rc_w) + diff_ind[k] * rc_w; /* mix(hemi, diff_ind, rc_weight) */
                float indirect = hemi * p.ao;
                float color = Lo[k] + indirect;
                color = color + p.emissive[k];
                o[k] = color;
            }
            o[3] = p.alpha;
        }
    }
    return pairs;
}

uint64_t or_shade(const GpuCameraUniforms* cam, const DeferredGlobals* g, const hl_gbuffer* gb,
                  const GpuLight* lights, uint32_t num_lights, const hl_light_grid* grid,
                  float* out, uint32_t out_pitch, uint32_t flags, uint32_t row0, uint32_t row1) {
    return or_shade_ex(cam, g, gb, lights, num_lights, grid, NULL, NULL, out, out_pitch, flags, row0, row1);
}

uint64_t or_shade_shadowed(const GpuCameraUniforms* cam, const DeferredGlobals* g, const hl_gbuffer* gb,
                           const GpuLight* lights, uint32_t num_lights, const hl_light_grid* grid,
                           const hl_shadow_maps* sh, float* out, uint32_t out_pitch, uint32_t flags,
                           uint32_t row0, uint32_t row1) {
    return or_shade_ex(cam, g, gb, lights, num_lights, grid, sh, NULL, out, out_pitch, flags, row0, row1);
}

uint64_t or_count_pairs(const GpuCameraUniforms* cam, const DeferredGlobals* g,
                        const hl_gbuffer* gb, const GpuLight* lights, uint32_t num_lights,
                        const hl_light_grid* grid) {
    DeferredGlobals g0 = *g;
    g0.debug_mode = 0;
    return or_shade(cam, &g0, gb, lights, num_lights, grid, NULL, 0, 0, 0, 0);
}

uint64_t or_check_conservative(const GpuCameraUniforms* cam, const DeferredGlobals* g,
                               const hl_gbuffer* gb, const GpuLight* lights, uint32_t num_lights,
                               const hl_light_grid* grid, uint32_t cap, uint32_t stride,
                               uint64_t* checked) {
    uint64_t bad = 0, seen = 0;
    if (stride == 0) stride = 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad, seen)
    for (int64_t yy = 0; yy < (int64_t)gb->height; yy += stride) {
        const uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < gb->width; x += stride) {
            const float depth = or_load_depth(&gb->depth, x, y);
            if (depth >= 1.0f) continue;
            or_px p;
            or_setup_pixel(cam, gb, x, y, depth, &p);
            const uint32_t* list;
            uint32_t cnt;
            or_cell_list(grid, g, x, y, depth, &list, &cnt);
            if (cap && cnt >= cap) continue; /* truncated cell: reference semantics drop lights */
            for (uint32_t i = 0; i < num_lights; ++i) {
                float c[3];
                or_pbr_direct_light(&lights[i], p.wp, p.N, p.V, p.F0, p.albedo, p.rough, p.metal, c);
                if (c[0] == 0.0f && c[1] == 0.0f && c[2] == 0.0f) continue;
                ++seen;
                int found = 0;
                for (uint32_t j = 0; j < cnt; ++j)
                    if (list[j] == i) { found = 1; break; }
                if (!found) ++bad;
            }
        }
    }
    if (checked) *checked = seen;
    return bad;
}
