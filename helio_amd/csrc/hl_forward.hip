// hl_forward.hip -- forward-lit fragment shading for gfx950: fs_main of the
// forward-lit pass (crates/passes/3d/helio-pass-forward-lit/shaders/
// forward_lit.wgsl:440-493) over a batch of fragments, the second consumer of
// the light grid the culls build.
//
//  k_forward_prep  one thread per light: its shading record, with the world
//                  position and direction of the light's Transform
//                  (forward_lit.wgsl:337-372, 468-477).  The shader evaluates
//                  the rotation (six sin/cos and two quaternion products) for
//                  every (fragment, light) pair; here it is done once per light.
//  k_forward_lit   one wavefront per 64 consecutive fragments (a block of 256):
//                  the fragment's set-up from its 80-byte record, then the
//                  deferred pass's list walk (distinct cells of the wave, lights
//                  refined against the AABB of the wave's fragments in each
//                  cell and staged in LDS) and pbr_direct_light without the
//                  radiance cut-off (forward_lit.wgsl:374-438).  Fragments of
//                  one screen region share a cell, so coherent fragment order
//                  keeps a wave on one or a few lists.
#include "hl_lighting.hpp"

namespace hl {

namespace {

// radians(d) of WGSL: d * (pi / 180)
constexpr float kDegToRad = 0.017453292519943295f;

// light_direction_from_rotation (forward_lit.wgsl:337-372): -Z rotated by
// q = qy(yaw) * qx(pitch) * qz(roll), then normalised; the quaternion product
// and rotation in the shader's order, one rounding per operation.
__device__ __forceinline__ float3 direction_from_rotation(float pitch_deg, float yaw_deg, float roll_deg) {
#pragma clang fp contract(off)
    const float hx = pitch_deg * kDegToRad * 0.5f, hy = yaw_deg * kDegToRad * 0.5f,
                hz = roll_deg * kDegToRad * 0.5f;
    const float sy = sinf(hy), cy = cosf(hy), sx = sinf(hx), cx = cosf(hx), sz = sinf(hz), cz = cosf(hz);
    // a = qy * qx with qy = (0, sy, 0, cy), qx = (sx, 0, 0, cx) (quat_mul, :317-324)
    const float a0 = cy * sx + 0.0f * cx + sy * 0.0f - 0.0f * 0.0f;
    const float a1 = cy * 0.0f - 0.0f * 0.0f + sy * cx + 0.0f * sx;
    const float a2 = cy * 0.0f + 0.0f * 0.0f - sy * sx + 0.0f * cx;
    const float a3 = cy * cx - 0.0f * sx - sy * 0.0f - 0.0f * 0.0f;
    // q = a * qz with qz = (0, 0, sz, cz)
    const float q0 = a3 * 0.0f + a0 * cz + a1 * sz - a2 * 0.0f;
    const float q1 = a3 * 0.0f - a0 * sz + a1 * cz + a2 * 0.0f;
    const float q2 = a3 * sz + a0 * 0.0f - a1 * 0.0f + a2 * cz;
    const float q3 = a3 * cz - a0 * 0.0f - a1 * 0.0f - a2 * sz;
    // quat_rotate_vec3 (:325-329) of v = (0, 0, -1): t = 2 (qv x v), v + w t + qv x t
    const float c0 = q1 * -1.0f - q2 * 0.0f, c1 = q2 * 0.0f - q0 * -1.0f, c2 = q0 * 0.0f - q1 * 0.0f;
    const float t0 = 2.0f * c0, t1 = 2.0f * c1, t2 = 2.0f * c2;
    const float e0 = q1 * t2 - q2 * t1, e1 = q2 * t0 - q0 * t2, e2 = q0 * t1 - q1 * t0;
    const float r0 = 0.0f + q3 * t0 + e0, r1 = 0.0f + q3 * t1 + e1, r2 = -1.0f + q3 * t2 + e2;
    return hl_x_normalize(r0, r1, r2);
}

__global__ __launch_bounds__(256) void k_forward_prep(const GpuLight* __restrict__ lights, uint32_t n,
                                                      const uint32_t* __restrict__ entity,
                                                      const hl_transform* __restrict__ transforms,
                                                      uint32_t num_transforms, HotLight* __restrict__ hot) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GpuLight* l = &lights[i];
    const float4 pr = *reinterpret_cast<const float4*>(l->position_range);
    const float4 dro = *reinterpret_cast<const float4*>(l->direction_outer);
    const float4 ci = *reinterpret_cast<const float4*>(l->color_intensity);
    const uint32_t type = l->light_type;
    float3 pos = make_float3(pr.x, pr.y, pr.z), dir = make_float3(dro.x, dro.y, dro.z);
    if (entity) {
        const uint32_t e = entity[i];
        if (e >= num_transforms) {
            // no transform: a record no cell AABB accepts (the light is skipped)
            const float inf = __builtin_inff();
            HotLight h;
            h.p0 = make_float4(inf, inf, inf, 0.0f);
            h.p1 = h.p2 = h.p3 = h.p4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            h.p2.w = kNoCone;
            h.p3.w = __uint_as_float(kNoShadowWord);
            hot[i] = h;
            return;
        }
        const float* t = transforms[e].position;
        pos = make_float3(t[0], t[1], t[2]);
        const float* r = transforms[e].rotation;
        dir = direction_from_rotation(r[0], r[1], r[2]);
    }
    // forward_lit.wgsl samples no shadow map
    hot[i] = make_hot_light(pos.x, pos.y, pos.z, dir.x, dir.y, dir.z, ci, pr.w, dro.w, l->inner_angle, type,
                            kNoShadowWord, false);
}

constexpr int kFwdModeShade = 0, kFwdModeCount = 1;

struct FwdSmem {
    HotLight light[4][kStage];
    float eye[3];
    float vp_z[4], vp_w[4];  // rows 2 and 3 of view_proj
    float split[HL_MAX_SLICES + 1];
};

template <int GRID, int MODE>
__global__ __launch_bounds__(256, 5) void k_forward_lit(ForwardArgs a) {
    __shared__ FwdSmem sm;
    const uint32_t tid = threadIdx.x;
    {
        const float* vp = a.cam->view_proj;
        if (tid < 3u) sm.eye[tid] = a.cam->position_near[tid];
        else if (tid >= 4u && tid < 8u) sm.vp_z[tid - 4u] = vp[(tid - 4u) * 4u + 2u];
        else if (tid >= 8u && tid < 12u) sm.vp_w[tid - 8u] = vp[(tid - 8u) * 4u + 3u];
        else if (GRID == HL_GRID_CLUSTERS && tid >= 64u && tid <= 64u + HL_MAX_SLICES)
            sm.split[tid - 64u] = a.depth_split[tid - 64u];
    }
    __syncthreads();
    const uint32_t lane = tid & 63u, wave = hl_uniform(tid >> 6);
    const uint32_t i = blockIdx.x * 256u + tid;
    const bool have = i < a.num_frags;
    float4 r0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), r1 = r0, r2 = r0, r3 = r0, r4 = r0;
    if (have) {
        const float4* fp = reinterpret_cast<const float4*>(a.frags + i);
        r0 = hl_ld4(fp);
        r1 = hl_ld4(fp + 1);
        r2 = hl_ld4(fp + 2);
        r3 = hl_ld4(fp + 3);
        r4 = hl_ld4(fp + 4);
    }
    float4* out = MODE == kFwdModeShade && have ? reinterpret_cast<float4*>(a.out) + i : nullptr;
    // forward_lit.wgsl:444: alpha <= 0.001 discards (NaN alpha does not)
    const bool keep = have && !(r1.w <= 0.001f);
    if (MODE == kFwdModeShade && have && !keep) hl_store_nt(out, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    if (hl_ballot(keep) == 0) return;

    // ---- fragment set-up (forward_lit.wgsl:446-458)
    Pixel p;
    p.wx = r0.x;
    p.wy = r0.y;
    p.wz = r0.z;
    const uint32_t pixel = __float_as_uint(r0.w);
    p.nx = r1.x;
    p.ny = r1.y;
    p.nz = r1.z;
    {
        // the shader's V, IEEE-exact, and NdV in its order: at grazing view a
        // last-bit error of V is a large relative error of G1(NdV), and the
        // near-peak GGX terms depend on the last bits of V
        const float3 v = hl_x_normalize_r(sm.eye[0] - p.wx, sm.eye[1] - p.wy, sm.eye[2] - p.wz);
        p.vx = v.x;
        p.vy = v.y;
        p.vz = v.z;
        p.ndv = fmaxf(hl_x_dot3(p.nx, p.ny, p.nz, p.vx, p.vy, p.vz), 0.0f);
    }
    p.alb[0] = r2.x;
    p.alb[1] = r2.y;
    p.alb[2] = r2.z;
    p.ao = r2.w;
    p.f0[0] = r3.x;
    p.f0[1] = r3.y;
    p.f0[2] = r3.z;
    p.rough = r3.w;
    p.metal = r4.w;
    p.one_m = 1.0f - p.metal;
    {
        const float rr = p.rough + 1.0f;
        p.kk = rr * rr * 0.125f;
        p.kk1 = p.kk + 0.0001f;
        p.ev = p.ndv * (1.0f - p.kk) + p.kk + 0.0001f;
        p.gv = p.ndv * hl_f_rcp(p.ev);
        const float ra = p.rough * p.rough;
        p.a2 = ra * ra;
        p.a2gv = p.a2 * p.gv;
#pragma unroll
        for (int c = 0; c < 3; ++c) p.kd[c] = p.one_m * p.alb[c] * kInvPi;
    }

    // ---- the fragment's cell (forward_lit.wgsl:463-466 for tiles)
    const uint32_t x = pixel & 0xffffu, y = pixel >> 16;
    uint32_t cell = 0;
    bool in_grid;
    if (GRID == HL_GRID_TILES) {
        const uint32_t tx = x >> 4, ty = y >> 4;
        in_grid = tx < a.grid_x && ty < a.grid_y;
        cell = ty * a.tile_stride_x + tx;
    } else {
        const uint32_t cx = hl_floor_div(x, a.cell_w), cy = hl_floor_div(y, a.cell_h);
        in_grid = cx < a.grid_x && cy < a.grid_y;
        float depth;
        {
            // NDC depth of the world position: rows 2 and 3 of view_proj
#pragma clang fp contract(off)
            float z = sm.vp_z[0] * p.wx;
            z = z + sm.vp_z[1] * p.wy;
            z = z + sm.vp_z[2] * p.wz;
            z = z + sm.vp_z[3];
            float w = sm.vp_w[0] * p.wx;
            w = w + sm.vp_w[1] * p.wy;
            w = w + sm.vp_w[2] * p.wz;
            w = w + sm.vp_w[3];
            depth = z / w;
        }
        // the largest slice k with depth_split[k] <= depth (0 if none); the
        // thresholds are non-decreasing
        uint32_t k = 0;
#pragma unroll
        for (uint32_t step = HL_MAX_SLICES / 2; step; step >>= 1) {
            const uint32_t j = k + step;
            if (j < a.grid_z && sm.split[j] <= depth) k = j;
        }
        cell = cx + a.grid_x * (cy + a.grid_y * k);
    }
    const bool walk = keep && in_grid;

    HotLight* s_l = sm.light[wave];
    if (MODE == kFwdModeCount) {
        uint32_t pairs = 0;
        walk_lights<GRID>(a, s_l, p, walk, cell, [&](const HotLight& L, uint32_t, bool mine) {
            if (!mine || __builtin_isinf(L.p0.w)) {
                pairs += mine ? 1u : 0u;
                return;
            }
            const float tx = L.p0.x - p.wx, ty = L.p0.y - p.wy, tz = L.p0.z - p.wz;
            pairs += range_test(L, tx, ty, tz, tx * tx + ty * ty + tz * tz) ? 1u : 0u;
        });
        // exact: at most 64 fragments x lights per cell per wave, far below 2^24
        const uint32_t tot = (uint32_t)dpp_sum((float)pairs);
        if (lane == 0 && tot) atomicAdd(a.pair_counter, (unsigned long long)tot);
        return;
    }
    float lo[3] = {0.0f, 0.0f, 0.0f};
    walk_lights<GRID>(a, s_l, p, walk, cell, [&](const HotLight& L, uint32_t, bool mine) {
        if (!mine) return;
        Term t;
        if (eval_term<false, false>(L, p, t)) {
            const float aw = t.atten * t.ndl;
            lo[0] += t.B[0] * (L.p1.x * aw);
            lo[1] += t.B[1] * (L.p1.y * aw);
            lo[2] += t.B[2] * (L.p1.z * aw);
        }
    });
    if (keep) {
        // forward_lit.wgsl:488-491: Lo + ambient + emissive, alpha
#pragma clang fp contract(off)
        const float o0 = lo[0] + a.amb[0] * p.alb[0] * p.ao + r4.x;
        const float o1 = lo[1] + a.amb[1] * p.alb[1] * p.ao + r4.y;
        const float o2 = lo[2] + a.amb[2] * p.alb[2] * p.ao + r4.z;
        hl_store_nt(out, make_float4(o0, o1, o2, r1.w));
    }
}

template <int GRID, int MODE>
hipError_t launch_fwd(hipStream_t s, const ForwardArgs& a) {
    const uint32_t blocks = (a.num_frags + 255u) / 256u;
    k_forward_lit<GRID, MODE><<<blocks, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_forward_prep(hipStream_t s, const GpuLight* lights, uint32_t n, const uint32_t* entity,
                               const hl_transform* transforms, uint32_t num_transforms, HotLight* hot) {
    if (n == 0) return hipSuccess;
    k_forward_prep<<<(n + 255u) / 256u, 256, 0, s>>>(lights, n, entity, transforms, num_transforms, hot);
    return hipGetLastError();
}

hipError_t launch_forward_lit(hipStream_t s, const ForwardArgs& a) {
    if (a.num_frags == 0) return hipSuccess;
    const bool count = a.pair_counter != nullptr;
    if (a.grid_mode == HL_GRID_TILES)
        return count ? launch_fwd<HL_GRID_TILES, kFwdModeCount>(s, a) : launch_fwd<HL_GRID_TILES, kFwdModeShade>(s, a);
    return count ? launch_fwd<HL_GRID_CLUSTERS, kFwdModeCount>(s, a)
                 : launch_fwd<HL_GRID_CLUSTERS, kFwdModeShade>(s, a);
}

}  // namespace hl
