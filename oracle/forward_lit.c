/*
 * forward_lit.c -- FP32 restatement of the forward-lit pass's fragment shader
 * (crates/passes/3d/helio-pass-forward-lit/shaders/forward_lit.wgsl), the
 * second consumer of the light grid.
 *
 * TEST INFRASTRUCTURE ONLY: used by the tests as the checker of
 * hl_forward_lit, never by the library.  Built with -ffp-contract=off: one
 * IEEE operation per source operation, in the orders fixed in oracle.h.
 * sin and cos are the C library's (the shader leaves them to the backend).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static const float FL_PI = 3.14159265359f; /* pbr_eval.wgsl:6 */

static float fl_dot(const float a[3], const float b[3]) {
    float s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}
static void fl_normalize(float v[3]) {
    const float l = sqrtf(fl_dot(v, v));
    v[0] = v[0] / l;
    v[1] = v[1] / l;
    v[2] = v[2] / l;
}
static float fl_max(float a, float b) { return a > b ? a : b; }
static float fl_min(float a, float b) { return a < b ? a : b; }
static float fl_clamp(float x, float lo, float hi) { return fl_min(fl_max(x, lo), hi); }

/* quat_mul, forward_lit.wgsl:317-324 (xyzw) */
static void fl_quat_mul(const float a[4], const float b[4], float r[4]) {
    r[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
    r[1] = a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0];
    r[2] = a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3];
    r[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
}

static void fl_cross(const float a[3], const float b[3], float r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* light_direction_from_rotation, forward_lit.wgsl:337-372: -Z rotated by
 * q = qy(yaw) * qx(pitch) * qz(roll), radians(d) = d * (pi / 180). */
void or_light_direction_from_rotation(const float rot_deg[3], float out[3]) {
    const float k = 0.017453292519943295f;
    const float pitch = rot_deg[0] * k, yaw = rot_deg[1] * k, roll = rot_deg[2] * k;
    const float hy = yaw * 0.5f, hx = pitch * 0.5f, hz = roll * 0.5f;
    const float qy[4] = {0.0f, sinf(hy), 0.0f, cosf(hy)};
    const float qx[4] = {sinf(hx), 0.0f, 0.0f, cosf(hx)};
    const float qz[4] = {0.0f, 0.0f, sinf(hz), cosf(hz)};
    float qyx[4], q[4];
    fl_quat_mul(qy, qx, qyx);
    fl_quat_mul(qyx, qz, q);
    /* quat_rotate_vec3, forward_lit.wgsl:325-329: v + w t + qv x t, t = 2 (qv x v) */
    const float v[3] = {0.0f, 0.0f, -1.0f};
    const float qv[3] = {q[0], q[1], q[2]};
    float c[3], t[3], c2[3];
    fl_cross(qv, v, c);
    t[0] = 2.0f * c[0];
    t[1] = 2.0f * c[1];
    t[2] = 2.0f * c[2];
    fl_cross(qv, t, c2);
    for (int i = 0; i < 3; ++i) out[i] = v[i] + q[3] * t[i] + c2[i];
    fl_normalize(out);
}

/* pbr_direct_light of the forward shader, forward_lit.wgsl:374-438 (isotropic,
 * no subsurface, sf = 1; no radiance cut-off). */
void or_forward_pbr_direct_light(const GpuLight* light, const float light_pos[3], const float light_dir[3],
                                 const float wp[3], const float N[3], const float V[3], const float F0[3],
                                 const float albedo[3], float roughness, float metallic, float out[3]) {
    float L[3], radiance[3];
    out[0] = out[1] = out[2] = 0.0f;
    if (light->light_type == 0u) {
        L[0] = -light_dir[0];
        L[1] = -light_dir[1];
        L[2] = -light_dir[2];
        fl_normalize(L);
        for (int k = 0; k < 3; ++k) radiance[k] = light->color_intensity[k] * light->color_intensity[3];
    } else {
        float to_light[3] = {light_pos[0] - wp[0], light_pos[1] - wp[1], light_pos[2] - wp[2]};
        const float dist = sqrtf(fl_dot(to_light, to_light));
        if (dist > light->position_range[3]) return;
        L[0] = to_light[0] / dist;
        L[1] = to_light[1] / dist;
        L[2] = to_light[2] / dist;
        float atten = 1.0f / (dist * dist + 0.0001f);
        const float nd = dist / light->position_range[3];
        atten = atten * fl_max(0.0f, 1.0f - nd * nd * nd * nd);
        if (light->light_type == 2u) {
            const float negL[3] = {-L[0], -L[1], -L[2]};
            const float cos_a = fl_dot(negL, light_dir);
            const float e0 = light->direction_outer[3], e1 = light->inner_angle;
            const float t = fl_clamp((cos_a - e0) / (e1 - e0), 0.0f, 1.0f);
            atten = atten * (t * t * (3.0f - 2.0f * t));
        }
        for (int k = 0; k < 3; ++k) radiance[k] = light->color_intensity[k] * light->color_intensity[3] * atten;
    }
    const float NdL = fl_max(fl_dot(N, L), 0.0f);
    if (NdL == 0.0f) return;
    float H[3] = {V[0] + L[0], V[1] + L[1], V[2] + L[2]};
    fl_normalize(H);
    const float HdV = fl_max(fl_dot(H, V), 0.0f);
    const float u = fl_clamp(1.0f - HdV, 0.0f, 1.0f);
    const float u2 = u * u;
    const float p5 = u2 * u2 * u;
    const float NdV = fl_max(fl_dot(N, V), 0.0f);
    const float a = roughness * roughness;
    const float a2 = a * a;
    const float NdH = fl_max(fl_dot(N, H), 0.0f);
    const float denom = NdH * NdH * (a2 - 1.0f) + 1.0f;
    const float D = a2 / (FL_PI * denom * denom + 0.0001f);
    const float rr = roughness + 1.0f;
    const float kk = (rr * rr) / 8.0f;
    const float G = NdV / (NdV * (1.0f - kk) + kk + 0.0001f) * (NdL / (NdL * (1.0f - kk) + kk + 0.0001f));
    const float sden = 4.0f * NdV * NdL + 0.0001f;
    for (int k = 0; k < 3; ++k) {
        const float F = F0[k] + (1.0f - F0[k]) * p5;
        const float kD = (1.0f - F) * (1.0f - metallic);
        const float spec = D * G * F / sden;
        const float diff = kD * albedo[k] / FL_PI;
        out[k] = (diff + spec) * radiance[k] * NdL;
    }
}

/* Light list of a fragment's pixel (forward_lit.wgsl:463-466 for tiles; the
 * slice of the NDC depth of the world position for clusters).  0 when the
 * pixel lies outside the grid. */
static uint32_t fl_cell_list(const GpuCameraUniforms* cam, const ForwardLitGlobals* g,
                             const hl_light_grid* grid, uint32_t pixel, const float wp[3],
                             const uint32_t** list) {
    const uint32_t x = pixel & 0xffffu, y = pixel >> 16;
    if (grid->mode == HL_GRID_TILES) {
        const uint32_t tx = x / HL_TILE_SIZE, ty = y / HL_TILE_SIZE;
        if (tx >= grid->num_tiles_x || ty >= grid->num_tiles_y) return 0;
        const uint32_t tile = ty * g->num_tiles_x + tx;
        *list = grid->lists_d + (size_t)tile * HL_MAX_LIGHTS_PER_TILE;
        return grid->counts_d[tile];
    }
    const hl_cluster_params* cp = grid->clusters;
    const uint32_t cx = x / cp->cell_w, cy = y / cp->cell_h;
    if (cx >= cp->grid_x || cy >= cp->grid_y) return 0;
    const float* m = cam->view_proj;
    float z = m[2] * wp[0];
    z = z + m[6] * wp[1];
    z = z + m[10] * wp[2];
    z = z + m[14];
    float w = m[3] * wp[0];
    w = w + m[7] * wp[1];
    w = w + m[11] * wp[2];
    w = w + m[15];
    const uint32_t cz = or_slice_of_depth(cp, z / w);
    const uint32_t c = cx + cp->grid_x * (cy + cp->grid_y * cz);
    *list = grid->lists_d + grid->offsets_d[c];
    return grid->counts_d[c];
}

/* fs_main of forward_lit.wgsl:440-493 for every fragment. */
uint64_t or_forward_lit(const GpuCameraUniforms* cam, const ForwardLitGlobals* g, const hl_fragment* frags,
                        uint32_t n, const GpuLight* lights, uint32_t num_lights, const uint32_t* entity,
                        const hl_transform* transforms, uint32_t num_transforms, const hl_light_grid* grid,
                        float* out) {
    uint64_t pairs = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : pairs)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const hl_fragment* f = &frags[i];
        float* o = out ? out + 4 * i : NULL;
        if (f->alpha <= 0.001f) { /* discard */
            if (o) o[0] = o[1] = o[2] = o[3] = 0.0f;
            continue;
        }
        float V[3] = {cam->position_near[0] - f->world_pos[0], cam->position_near[1] - f->world_pos[1],
                      cam->position_near[2] - f->world_pos[2]};
        fl_normalize(V);
        float Lo[3] = {0.0f, 0.0f, 0.0f};
        const uint32_t* list = NULL;
        const uint32_t cnt = fl_cell_list(cam, g, grid, f->pixel, f->world_pos, &list);
        for (uint32_t j = 0; j < cnt; ++j) {
            const uint32_t li = list[j];
            if (li >= num_lights) continue;
            const GpuLight* l = &lights[li];
            float pos[3], dir[3];
            if (entity) {
                const uint32_t e = entity[li];
                if (e >= num_transforms) continue;
                memcpy(pos, transforms[e].position, sizeof pos);
                or_light_direction_from_rotation(transforms[e].rotation, dir);
            } else {
                memcpy(pos, l->position_range, sizeof pos);
                memcpy(dir, l->direction_outer, sizeof dir);
            }
            if (l->light_type != 0u) { /* forward_lit.wgsl:478-481 */
                const float d[3] = {pos[0] - f->world_pos[0], pos[1] - f->world_pos[1], pos[2] - f->world_pos[2]};
                if (sqrtf(fl_dot(d, d)) > l->position_range[3]) continue;
            }
            ++pairs;
            if (!o) continue;
            float c[3];
            or_forward_pbr_direct_light(l, pos, dir, f->world_pos, f->normal, V, f->specular_f0, f->albedo,
                                        f->roughness, f->metallic, c);
            Lo[0] = Lo[0] + c[0];
            Lo[1] = Lo[1] + c[1];
            Lo[2] = Lo[2] + c[2];
        }
        if (!o) continue;
        for (int k = 0; k < 3; ++k) { /* forward_lit.wgsl:488-490 */
            const float ambient = g->ambient_color[k] * g->ambient_intensity * f->albedo[k] * f->ao;
            o[k] = Lo[k] + ambient + f->emissive[k];
        }
        o[3] = f->alpha;
    }
    return pairs;
}
