// hl_shade.hip -- deferred shading (fs_main subset) and its adjoint for gfx950.
//
// Launch shape: 256-thread blocks covering 32x8 pixels; each wavefront owns an
// 8x8 pixel tile.  A wavefront
//   1. loads depth and the G-buffer (rows of 32 px per block), reconstructs the
//      world position exactly and finds each pixel's cell (tile or cluster);
//   2. for each distinct cell among its pixels: reduces the AABB of those
//      pixels' world positions (DPP), streams the cell's list in chunks of 64,
//      tests each light sphere against the AABB and compacts the survivors --
//      index and 64-byte light record -- into LDS with ballot + prefix popcount
//      (list order preserved);
//   3. walks the LDS list (uniform address, broadcast reads) and evaluates
//      pbr_direct_light for its pixels.
// Every pixel therefore sums exactly the lights of its own cell list that pass
// the range test, in list order, as deferred_lighting.wgsl:1125-1145 does.
// The adjoint re-derives the same per-wave lists instead of storing them.
#include "hl_lighting.hpp"
#include "hl_rc.hpp"
#include "hl_shadow.hpp"

namespace hl {

// Reverse-mode derivative of one active term given dL/dout (gl).  Per-pixel
// gradients accumulate into dA (albedo), dR (roughness), dM (metallic); the
// light's position and intensity gradients are returned in g (x, y, z, I).
__device__ __forceinline__ void adjoint_term(const HotLight& L, const Pixel& p, const Term& t,
                                             const float gl[3], float dA[3], float& dR, float& dM,
                                             float4& g) {
    const float m = p.metal, r = p.rough;
    const float inv_den2 = hl_f_rcp(t.den2);
    const float DG_den2 = t.D * t.G * inv_den2;
    const float invQ = hl_f_rcp(t.Q);
    // dD/dr and dG/dr (pbr_eval.wgsl:12-43)
    const float dDdr = (invQ - t.a2 * 2.0f * kPi * t.den * t.ndh * t.ndh * invQ * invQ) * 4.0f * r * r * r;
    const float inv_ev = hl_f_rcp(p.ev), inv_el = hl_f_rcp(t.el);
    const float dgdk_v = -p.ndv * (1.0f - p.ndv) * inv_ev * inv_ev;
    const float dgdk_l = -t.ndl * (1.0f - t.ndl) * inv_el * inv_el;
    const float dGdr = (dgdk_v * t.gl + p.gv * dgdk_l) * (r + 1.0f) * 0.25f;
    const float dSdr = (dDdr * t.G + t.D * dGdr) * inv_den2;
    const float atten = t.att0 * t.fall * t.s;
    float adj_NdL = 0.0f, adj_atten = 0.0f, adj_D = 0.0f, adj_G = 0.0f, adj_den2 = 0.0f, adj_P = 0.0f;
    float dI = 0.0f;
    // color = (color * intensity) / intensity: an active term has radiance
    // >= 0.002, hence intensity != 0
    const float inv_I = hl_f_rcp(L.p3.y);
    const float cI[3] = {L.p1.x, L.p1.y, L.p1.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float W = t.rad[c] * t.ndl;
        const float gw = gl[c] * W;
        const float dBdF = -(1.0f - m) * p.alb[c] * kInvPi + DG_den2;
        const float kD = (1.0f - t.F[c]) * p.one_m;
        dA[c] += gw * (kD * kInvPi + (p.f0_free[c] ? m * (1.0f - t.P) * dBdF : 0.0f));
        dM += gw * (-(1.0f - t.F[c]) * p.alb[c] * kInvPi +
                    (p.f0_free[c] ? (p.alb[c] - 0.04f) * (1.0f - t.P) * dBdF : 0.0f));
        dR += gw * dSdr * t.F[c];
        dI += gl[c] * t.B[c] * cI[c] * inv_I * atten * t.ndl;
        const float adj_B = gw;
        adj_NdL += gl[c] * t.B[c] * t.rad[c];
        adj_atten += gl[c] * t.B[c] * t.ndl * cI[c];
        adj_D += adj_B * t.G * t.F[c] * inv_den2;
        adj_G += adj_B * t.D * t.F[c] * inv_den2;
        adj_den2 -= adj_B * DG_den2 * t.F[c] * inv_den2;
        adj_P += adj_B * dBdF * (1.0f - p.f0[c]);
    }
    g = make_float4(0.0f, 0.0f, 0.0f, dI);
    if (__builtin_isinf(L.p0.w)) return;  // directional: no position
    float ahx = 0.0f, ahy = 0.0f, ahz = 0.0f;
    if (t.hdv > 0.0f && t.u > 0.0f) {
        const float u2 = t.u * t.u;
        const float a = adj_P * (-5.0f * u2 * u2);
        ahx += a * p.vx;
        ahy += a * p.vy;
        ahz += a * p.vz;
    }
    if (t.ndh > 0.0f) {
        const float dDdNdH = -t.a2 * invQ * invQ * (2.0f * kPi * t.den) * (2.0f * t.ndh * (t.a2 - 1.0f));
        const float a = adj_D * dDdNdH;
        ahx += a * p.nx;
        ahy += a * p.ny;
        ahz += a * p.nz;
    }
    adj_NdL += adj_G * p.gv * (p.kk + 0.0001f) * inv_el * inv_el;
    adj_NdL += adj_den2 * 4.0f * p.ndv;
    const float hd = t.hx * ahx + t.hy * ahy + t.hz * ahz;
    float alx = (ahx - t.hx * hd) * t.inv_h + adj_NdL * p.nx;
    float aly = (ahy - t.hy * hd) * t.inv_h + adj_NdL * p.ny;
    float alz = (ahz - t.hz * hd) * t.inv_h + adj_NdL * p.nz;
    const float adj_att0 = adj_atten * t.fall * t.s;
    const float adj_fall = adj_atten * t.att0 * t.s;
    const float adj_s = adj_atten * t.att0 * t.fall;
    const float R = L.p0.w;
    const float invR = hl_f_rcp(R);
    float adj_d = adj_att0 * (-2.0f * t.d * t.att0 * t.att0);
    if (t.fall > 0.0f) {
        const float dr = t.d * invR;
        adj_d += adj_fall * (-4.0f * dr * dr * dr * invR);
    }
    if (L.p2.w != kNoCone && t.dsdcos != 0.0f) {
        const float a = adj_s * t.dsdcos;
        alx -= a * L.p2.x;
        aly -= a * L.p2.y;
        alz -= a * L.p2.z;
    }
    const float ld = t.lx * alx + t.ly * aly + t.lz * alz;
    g.x = (alx - t.lx * ld) * t.inv_d + adj_d * t.lx;
    g.y = (aly - t.ly * ld) * t.inv_d + adj_d * t.ly;
    g.z = (alz - t.lz * ld) * t.inv_d + adj_d * t.lz;
}

// -------------------------------------------------------------- block tile
// Shades block tile b of a view: 4 wavefronts side by side, each owning a
// TW x (64/TW) pixel tile (TW = 1 << a.tw_log2: 8x8, 16x4 or 32x2).
// Per-view constants of the pixel set-up, staged in LDS once per view rather
// than read per tile: the camera terms and the slice thresholds.
struct ViewConst {
    float ivp[16];                   // inv_view_proj
    float eye[3];                    // position_near.xyz
    float p22, p32;                  // proj[10], proj[14]
    float split[HL_MAX_SLICES + 1];  // depth_split
};

struct ShadeSmem {
    HotLight light[4][kStage];
    ViewConst vc;
};
// LDS of the adjoint modes only (declared by their kernels alone, so that the
// forward's occupancy is not charged for it)
struct AdjSmem {
    KeptLights kept[4];      // fused: the lights each wave's walk visited
    float gl[4][3][64];      // fused: the waves' partial radiance sums
    float red[1][4][5][64];  // the waves' partial per-pixel gradient sums
};

// Fills sm.vc from the view's arguments; the caller synchronises the block
// before the first tile reads it.
__device__ __forceinline__ void load_view_const(const ShadeArgs& a, ShadeSmem& sm) {
    const GpuCameraUniforms* cam = a.cam;
    const uint32_t t = threadIdx.x;
    if (t < 16u) sm.vc.ivp[t] = cam->inv_view_proj[t];
    else if (t < 19u) sm.vc.eye[t - 16u] = cam->position_near[t - 16u];
    else if (t == 19u) sm.vc.p22 = cam->proj[10];
    else if (t == 20u) sm.vc.p32 = cam->proj[14];
    else if (t >= 64u && t <= 64u + HL_MAX_SLICES) sm.vc.split[t - 64u] = a.depth_split[t - 64u];
}

// F16: the output target is RGBA16F (a compile-time choice, so that the
// RGBA32F kernels hold no code of the other format)
// The per-pixel set-up of one lane (deferred_lighting.wgsl:999-1086): the
// world position and view vector from the depth, the pixel's cell, and the
// decoded G-buffer texels with the material terms.  A and E are the decoded
// albedo and emissive texels.  Lanes without a pixel (valid false) get
// harmless values.
template <int FA, int FN, int GRID, int MODE>
__device__ __forceinline__ void setup_pixel(const ShadeArgs& a, const ViewConst& vc, uint32_t x, uint32_t y,
                                            float depth, bool valid, Pixel& p, uint32_t& cell, float4& A,
                                            float4& E) {
    // the four G-buffer fetches go out together; the set-up that needs the
    // depth only (world position, view vector, cell) runs while they are in
    // flight, and the texels are decoded after it
    typename TexRaw<FA>::type rA{}, rO{};
    typename TexRaw<FN>::type rN{}, rE{};
    if (valid) {
        rA = hl_fetch<FA>(a.albedo, x, y);
        rN = hl_fetch<FN>(a.normal, x, y);
        rO = hl_fetch<FA>(a.orm, x, y);
        rE = hl_fetch<FN>(a.emissive, x, y);
    }
    __builtin_amdgcn_sched_barrier(0);
    {
        // world position exactly as deferred_lighting.wgsl:1003-1008 (the
        // range, cone and threshold decisions depend on it): uv = (px + 0.5) /
        // dims and w.xyz / w.w as IEEE quotients, from a reciprocal and two
        // fma corrections (hl_x_div_r) where the operands allow it
#pragma clang fp contract(off)
        const float u = hl_x_div_r((float)x + 0.5f, (float)a.width, a.rcp_w);
        const float v = hl_x_div_r((float)y + 0.5f, (float)a.height, a.rcp_h);
        const float ndx = u * 2.0f - 1.0f, ndy = 1.0f - v * 2.0f;
        const float* m = vc.ivp;
        const float hw = hl_x_mat_row(m, 3, ndx, ndy, depth, 1.0f);
        const float n0 = hl_x_mat_row(m, 0, ndx, ndy, depth, 1.0f);
        const float n1 = hl_x_mat_row(m, 1, ndx, ndy, depth, 1.0f);
        const float n2 = hl_x_mat_row(m, 2, ndx, ndy, depth, 1.0f);
        if (hl_div_r_ok(hw, n0, n1, n2) || !valid) {  // a lane without a pixel: any value
            const float y1 = hl_rcp_refined(hw);
            p.wx = hl_x_div_r(n0, hw, y1);
            p.wy = hl_x_div_r(n1, hw, y1);
            p.wz = hl_x_div_r(n2, hw, y1);
        } else {
            hl_rare();
            p.wx = n0 / hw;
            p.wy = n1 / hw;
            p.wz = n2 / hw;
        }
    }
    {
        // the shader's V, IEEE-exact (as N below): NdV at grazing view and the
        // near-peak GGX terms depend on its last bits
        const float3 v = hl_x_normalize_r(vc.eye[0] - p.wx, vc.eye[1] - p.wy, vc.eye[2] - p.wz);
        p.vx = v.x;
        p.vy = v.y;
        p.vz = v.z;
    }

    // ---- cell of this pixel
    cell = 0;
    if (GRID == HL_GRID_TILES) {
        cell = (y >> 4) * a.tile_stride_x + (x >> 4);
    } else {
        const uint32_t cx = hl_floor_div(x, a.cell_w), cy = hl_floor_div(y, a.cell_h);
        uint32_t cz = 0;
        if (valid && a.grid_z > 1u) {
            // guess from the view distance, then settle on the exact thresholds
            const float dist = vc.p32 * hl_f_rcp(depth + vc.p22);
            const float g = __builtin_amdgcn_logf(fmaxf(dist, 1e-30f)) * a.log_scale + a.log_bias;
            const int kmax = (int)a.grid_z - 1;
            int k = (int)floorf(g);
            k = k < 0 ? 0 : (k > kmax ? kmax : k);
            // the guess is almost always right: both of its thresholds at once
            const float* split = vc.split;
            const float lo = split[k], hi = split[k + 1];
            if (!((k == 0 || lo <= depth) && (k == kmax || hi > depth))) {
                hl_rare();
                while (k > 0 && split[k] > depth) --k;
                while (k < kmax && split[k + 1] <= depth) ++k;
            }
            cz = (uint32_t)k;
        }
        cell = cx + a.grid_x * (cy + a.grid_y * cz);
    }

    // zero bits decode to 0 in every format; lanes without a pixel keep N = +y
    A = hl_decode<FA>(rA);
    E = hl_decode<FN>(rE);
    const float4 O = hl_decode<FA>(rO);
    float4 Nr = hl_decode<FN>(rN);
    if (!valid) Nr = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
    {
        // the shader's N, IEEE-exact (the near-peak GGX terms and
        // shadow_factor's normal offset rely on it); NdV in the shader's order
        // (at grazing view NdV is a difference of terms near 1, and G1(NdV)
        // would carry a last-bit error of N or V as a relative error of
        // 1e-7 / NdV)
        const float3 n = hl_x_normalize_r(Nr.x, Nr.y, Nr.z);
        p.nx = n.x;
        p.ny = n.y;
        p.nz = n.z;
        p.ndv = fmaxf(hl_x_dot3(p.nx, p.ny, p.nz, p.vx, p.vy, p.vz), 0.0f);
    }
    p.alb[0] = A.x;
    p.alb[1] = A.y;
    p.alb[2] = A.z;
    p.ao = O.x;
    p.rough = O.y;
    p.metal = O.z;
    p.one_m = 1.0f - p.metal;
    if (MODE == kModeBackward || MODE == kModeFusedL2) {
        // the differentiated forward resolves F0 from albedo and metallic
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float v = 0.04f * (1.0f - p.metal) + p.alb[c] * p.metal;
            p.f0_free[c] = v > 0.0f && v < 0.999f;
            p.f0[c] = fminf(fmaxf(v, 0.0f), 0.999f);
        }
    } else {
        p.f0[0] = fminf(fmaxf(Nr.w, 0.0f), 0.999f);
        p.f0[1] = fminf(fmaxf(O.w, 0.0f), 0.999f);
        p.f0[2] = fminf(fmaxf(E.w, 0.0f), 0.999f);
    }
    {
        const float r1 = p.rough + 1.0f;
        p.kk = r1 * r1 * 0.125f;
        p.kk1 = p.kk + 0.0001f;
        p.ev = p.ndv * (1.0f - p.kk) + p.kk + 0.0001f;
        p.gv = p.ndv * hl_f_rcp(p.ev);
        const float ra = p.rough * p.rough;
        p.a2 = ra * ra;
        p.a2gv = p.a2 * p.gv;
#pragma unroll
        for (int c = 0; c < 3; ++c) p.kd[c] = p.one_m * p.alb[c] * kInvPi;
    }

}

// Ambient + emissive (deferred_lighting.wgsl:1196-1231), added to the sum of
// the lights at the end; mixv is the sky hemisphere mix.
__device__ __forceinline__ void ambient_terms(const ShadeArgs& a, const Pixel& p, const float4& E, float mixv[3],
                                              float amb[3]) {
    const float ht = p.ny * 0.5f + 0.5f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float sky = a.sky[c];
        mixv[c] = sky * 0.15f * (1.0f - ht) + sky * ht;
        amb[c] = mixv[c] * p.alb[c] * p.ao + (c == 0 ? E.x : (c == 1 ? E.y : E.z));
    }
}

// One light's adjoint for the pixels of the lanes with mine set: the
// per-pixel gradients go to dA, dR, dM; the light's position and intensity
// gradients, summed over the wave, to its slot copy `slot` (FP64 atomics by
// lanes 0..3).  gl is dL/d(radiance) of the lane's pixel.
__device__ __forceinline__ void adjoint_light(const ShadeArgs& a, const HotLight& L, uint32_t li, bool mine,
                                              const Pixel& p, const float gl[3], float dA[3], float& dR, float& dM,
                                              uint32_t slot) {
    const uint32_t lane = hl_lane();
    float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool act = false;
    if (mine) {
        Term t;
        act = eval_term<true>(L, p, t);
        if (act) adjoint_term(L, p, t, gl, dA, dR, dM, g);
    }
    if (hl_ballot(act) == 0) return;
    const float sx = dpp_sum(g.x), sy = dpp_sum(g.y), sz = dpp_sum(g.z), si = dpp_sum(g.w);
    if (lane < 4u) {
        const float v = lane == 0u ? sx : (lane == 1u ? sy : (lane == 2u ? sz : si));
        atomicAdd(a.d_light + 4 * ((size_t)slot * a.num_lights + li) + lane, (double)v);
    }
}

// Every mode but the fused forward + adjoint (fused_tile).
template <int FA, int FN, int GRID, int MODE, bool F16 = false>
__device__ __forceinline__ void shade_tile(const ShadeArgs& a, uint32_t b, ShadeSmem& sm, AdjSmem* as) {
    static_assert(MODE != kModeFusedL2, "fused_tile");
    const ViewConst& vc = sm.vc;
    b = hl_uniform(b);  // the item is the block's: its tile arithmetic stays scalar
    const uint32_t lane = threadIdx.x & 63u, wave = hl_uniform(threadIdx.x >> 6);
    constexpr bool kAdj = MODE == kModeBackward;
    // The backward mode splits a wave tile's adjoint over the block's four
    // waves (every fourth light each; the items are wave tiles, four per block
    // tile): its per-light work is several times the forward's, and the
    // heaviest tiles would otherwise set the launch's length.  The other modes
    // give each wave its own tile of the block tile.
    constexpr bool kSplit = kAdj;
    const uint32_t twl = a.tw_log2, tw = 1u << twl;
    const uint32_t th = 64u >> twl;
    const uint32_t bt = kSplit ? b >> 2 : b, sub = kSplit ? (b & 3u) : wave;
    const uint32_t bx = bt % a.block_tiles_x, by = bt / a.block_tiles_x;
    const uint32_t x = bx * (4u * tw) + sub * tw + (lane & (tw - 1u));
    const uint32_t y = by * th + (lane >> twl);
    const bool inside = x < a.width && y < a.height;
    const float depth = inside ? hl_load_depth(a.depth, x, y) : 1.0f;
    const bool valid = inside && depth < 1.0f;
    const bool lead = !kSplit || wave == 0u;  // writes the tile's shared results
    constexpr bool kRC = MODE == kModeExtended;
    constexpr bool kExt = MODE == kModeShadowed || kRC;  // exact N, shadow_factor
    constexpr bool kFwd = MODE == kModeForward || kExt;
    constexpr bool kWrites = kFwd;
    // the pixel's address, computed once (8 B per pixel for RGBA16F, else 16)
    constexpr bool out_f16 = F16;
    uint8_t* out_px = (kWrites && inside) ? reinterpret_cast<uint8_t*>(a.out) + (size_t)y * a.out_pitch +
                                                (size_t)x * (out_f16 ? 8u : 16u)
                                          : nullptr;
    const size_t plane = (size_t)a.width * a.height;
    const size_t pix = (size_t)y * a.width + x;
    if (inside && !valid && lead) {
        if (kWrites && (a.flags & HL_SHADE_CLEAR_SKY))
            hl_store_px(out_px, out_f16, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        if (kAdj)
            for (int k = 0; k < HL_PIXEL_GRAD_PLANES; ++k) hl_g(a.d_pixel)[k * plane + pix] = 0.0f;
    }
    if (hl_ballot(valid) == 0) return;

    // ---- per-pixel set-up (deferred_lighting.wgsl:999-1086)
    Pixel p;
    uint32_t cell;
    float4 A, E;
    setup_pixel<FA, FN, GRID, MODE>(a, vc, x, y, depth, valid, p, cell, A, E);

    // ---- debug views (deferred_lighting.wgsl:1032-1038)
    if (kFwd && a.debug_mode != 0u) {
        if (valid) {
            float4 o;
            if (a.debug_mode == 5u)
                o = make_float4(p.nx * 0.5f + 0.5f, p.ny * 0.5f + 0.5f, p.nz * 0.5f + 0.5f, 1.0f);
            else
                o = A;
            hl_store_px(out_px, out_f16, o);
        }
        return;
    }

    HotLight* s_l = sm.light[wave];
    float lo[3] = {0.0f, 0.0f, 0.0f};

    if (MODE == kModeCount) {
        uint32_t pairs = 0;
        uint32_t visits = 0;  // (wave, light) visits of the walk: uniform
        walk_lights<GRID>(a, s_l, p, valid, cell,
                          [&](const HotLight& L, uint32_t, bool mine) {
                              ++visits;
                              if (!mine || __builtin_isinf(L.p0.w)) {
                                  pairs += mine ? 1u : 0u;
                                  return;
                              }
                              const float tx = L.p0.x - p.wx, ty = L.p0.y - p.wy, tz = L.p0.z - p.wz;
                              const bool in = range_test(L, tx, ty, tz, tx * tx + ty * ty + tz * tz);
                              pairs += in ? 1u : 0u;
                          });
        // exact: at most 64 pixels x lights per cell per wave, far below 2^24
        const uint32_t tot = (uint32_t)dpp_sum((float)pairs);
        if (lane == 0 && tot) atomicAdd(a.pair_counter, (unsigned long long)tot);
        // the visits go to the counter two words on (a.pair_counter[2])
        if (lane == 0 && visits) atomicAdd(a.pair_counter + 2, (unsigned long long)visits);
        return;
    }

    // ambient + emissive, added to the sum of the lights at the end
    float mixv[3], amb[3], o[3];
    ambient_terms(a, p, E, mixv, amb);
    if (kRC && a.rc_fmt != 0u && valid) {
        // radiance-cascade indirect diffuse (deferred_lighting.wgsl:1159-1163,
        // 1189-1205): mix(hemisphere, kD_ibl rc_irr albedo, clamp(4 |rc_irr|, 0, 1))
        const float4 r4 = hl_ld4(a.rc_irr + pix);  // k_rc_irradiance, same N and position
        const float3 rc = make_float3(r4.x, r4.y, r4.z);
        const float w = fminf(fmaxf(sqrtf(rc.x * rc.x + rc.y * rc.y + rc.z * rc.z) * 4.0f, 0.0f), 1.0f);
        const float u = fminf(fmaxf(1.0f - p.ndv, 0.0f), 1.0f);
        const float u2 = u * u;
        const float p5 = u2 * u2 * u;
        const float rcv[3] = {rc.x, rc.y, rc.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // fresnel_schlick_roughness (pbr_eval.wgsl:63-66)
            const float f = p.f0[c] + (fmaxf(1.0f - p.rough, p.f0[c]) - p.f0[c]) * p5;
            const float diff = (1.0f - f) * (1.0f - p.metal) * rcv[c] * p.alb[c];
            const float hemi = mixv[c] * p.alb[c];
            amb[c] = (hemi * (1.0f - w) + diff * w) * p.ao + (c == 0 ? E.x : (c == 1 ? E.y : E.z));
        }
    }
    if (MODE != kModeBackward) {
        walk_lights<GRID>(a, s_l, p, valid, cell, [&](const HotLight& L, uint32_t, bool mine) {
            if (!mine) return;
            Term t;
            if (eval_term<false>(L, p, t)) {
                // radiance x NdL = p1 (atten NdL)
                float aw = t.atten * t.ndl;
                if (kExt) {
                    // pbr_direct_light's sf (deferred_lighting.wgsl:1136-1143); the
                    // word is the light's, so the branch is uniform
                    const uint32_t w = __float_as_uint(L.p3.w);
                    if (w != kNoShadowWord && a.shadow.layers != 0u)
                        aw *= shadow_factor(a.shadow, L, w >> 28, w & 0x0FFFFFFFu, p.wx, p.wy, p.wz, p.nx, p.ny,
                                            p.nz, (float)x + 0.5f, (float)y + 0.5f, vc.eye);
                }
                lo[0] += t.B[0] * (L.p1.x * aw);
                lo[1] += t.B[1] * (L.p1.y * aw);
                lo[2] += t.B[2] * (L.p1.z * aw);
            }
        });
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = lo[c] + amb[c];
    if (kFwd) {
        if (valid) hl_store_px(out_px, out_f16, make_float4(o[0], o[1], o[2], A.w));
        return;
    }

    // ---- adjoint (backward: dL/d(radiance) given)
    float gl[3] = {0.0f, 0.0f, 0.0f};
    if (valid) {
        const float4 g = hl_ld4((const float4*)(reinterpret_cast<const uint8_t*>(a.dout) + (size_t)y * a.dout_pitch) + x);
        gl[0] = g.x;
        gl[1] = g.y;
        gl[2] = g.z;
    }
    float dA[3] = {0.0f, 0.0f, 0.0f}, dR = 0.0f, dM = 0.0f;
    // per-light gradients go to one of kGradSlots copies (by wave) so that the
    // FP64 atomics of many waves on one light do not serialise on one address
    const uint32_t slot = (b * 4u + wave) % kGradSlots;
    uint32_t ord = 0;
    walk_lights<GRID>(a, s_l, p, valid, cell, [&](const HotLight& L, uint32_t li, bool mine) {
        if ((ord++ & 3u) != wave) return;
        adjoint_light(a, L, li, mine, p, gl, dA, dR, dM, slot);
    });
    // the waves' partial per-pixel gradients, added in wave order
    float (*red)[5][64] = as->red[0];
    red[wave][0][lane] = dA[0];
    red[wave][1][lane] = dA[1];
    red[wave][2][lane] = dA[2];
    red[wave][3][lane] = dR;
    red[wave][4][lane] = dM;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 3; ++c) dA[c] = ((red[0][c][lane] + red[1][c][lane]) + red[2][c][lane]) + red[3][c][lane];
    dR = ((red[0][3][lane] + red[1][3][lane]) + red[2][3][lane]) + red[3][3][lane];
    dM = ((red[0][4][lane] + red[1][4][lane]) + red[2][4][lane]) + red[3][4][lane];
    __syncthreads();
    if (valid && lead) {
#pragma unroll
        for (int c = 0; c < 3; ++c) dA[c] += gl[c] * mixv[c] * p.ao;
        const hl_gptr<float> dp = hl_g(a.d_pixel);
        dp[0 * plane + pix] = dA[0];
        dp[1 * plane + pix] = dA[1];
        dp[2 * plane + pix] = dA[2];
        dp[3 * plane + pix] = dR;
        dp[4 * plane + pix] = dM;
    }
}

// The fused forward + L2 adjoint of wave tile b & 3 of block tile b >> 2
// (the items are wave tiles, four per block tile): radiance into a.out, the
// loss sum((out - target)^2) into a.loss, and its gradients.  The block's
// four waves share the tile: each walks its lights (keeping the ones it
// visits in LDS) and evaluates every fourth one (ordinals = wave mod 4) in
// the forward and the adjoint alike; the waves' partial sums are added in
// wave order, so the output depends on the tile shape at the last bits.
template <int FA, int FN, int GRID>
__device__ __forceinline__ void fused_tile_split(const ShadeArgs& a, uint32_t b, ShadeSmem& sm, AdjSmem& as) {
    constexpr int MODE = kModeFusedL2;
    const ViewConst& vc = sm.vc;
    b = hl_uniform(b);
    const uint32_t lane = threadIdx.x & 63u, wave = hl_uniform(threadIdx.x >> 6);
    const uint32_t twl = a.tw_log2, tw = 1u << twl, th = 64u >> twl;
    const uint32_t bt = b >> 2, sub = b & 3u;
    const uint32_t bx = bt % a.block_tiles_x, by = bt / a.block_tiles_x;
    const uint32_t x = bx * (4u * tw) + sub * tw + (lane & (tw - 1u));
    const uint32_t y = by * th + (lane >> twl);
    const bool inside = x < a.width && y < a.height;
    const float depth = inside ? hl_load_depth(a.depth, x, y) : 1.0f;
    const bool valid = inside && depth < 1.0f;
    const bool lead = wave == 0u;  // writes the tile's results
    const size_t plane = (size_t)a.width * a.height;
    const size_t pix = (size_t)y * a.width + x;
    float4* out_px = inside ? (float4*)(reinterpret_cast<uint8_t*>(a.out) + (size_t)y * a.out_pitch) + x : nullptr;
    if (inside && !valid && lead) {
        if (a.flags & HL_SHADE_CLEAR_SKY) hl_store_nt(out_px, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        for (int k = 0; k < HL_PIXEL_GRAD_PLANES; ++k) hl_g(a.d_pixel)[k * plane + pix] = 0.0f;
    }
    if (hl_ballot(valid) == 0) return;  // uniform over the block: one tile

    Pixel p;
    uint32_t cell;
    float4 A, E;
    setup_pixel<FA, FN, GRID, MODE>(a, vc, x, y, depth, valid, p, cell, A, E);
    float mixv[3], amb[3];
    ambient_terms(a, p, E, mixv, amb);

    // ---- forward: every fourth light, the partial sums added in wave order
    HotLight* s_l = sm.light[wave];
    KeptLights& kept = as.kept[wave];
    KeptState ks;
    float lo[3] = {0.0f, 0.0f, 0.0f};
    uint32_t ord = 0;
    walk_lights<GRID, true>(a, s_l, p, valid, cell, [&](const HotLight& L, uint32_t, bool mine) {
        if ((ord++ & 3u) != wave || !mine) return;
        Term t;
        if (eval_term<false>(L, p, t)) {
            const float aw = t.atten * t.ndl;
            lo[0] += t.B[0] * (L.p1.x * aw);
            lo[1] += t.B[1] * (L.p1.y * aw);
            lo[2] += t.B[2] * (L.p1.z * aw);
        }
    }, &kept, &ks);
#pragma unroll
    for (int c = 0; c < 3; ++c) as.gl[wave][c][lane] = lo[c];
    __syncthreads();
    float o[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        o[c] = (((as.gl[0][c][lane] + as.gl[1][c][lane]) + as.gl[2][c][lane]) + as.gl[3][c][lane]) + amb[c];

    float gl[3] = {0.0f, 0.0f, 0.0f};
    float l2 = 0.0f;
    if (valid) {
        if (lead) hl_store_nt(out_px, make_float4(o[0], o[1], o[2], A.w));
        const float4 tg = hl_ld4((const float4*)(reinterpret_cast<const uint8_t*>(a.dout) + (size_t)y * a.dout_pitch) + x);
        const float e0 = o[0] - tg.x, e1 = o[1] - tg.y, e2 = o[2] - tg.z;
        gl[0] = 2.0f * e0;
        gl[1] = 2.0f * e1;
        gl[2] = 2.0f * e2;
        l2 = e0 * e0 + e1 * e1 + e2 * e2;
    }
    const float wl = dpp_sum(l2);
    if (lane == 0 && lead) atomicAdd(a.loss, (double)wl);

    // ---- adjoint: the same lights per wave, from LDS when they were kept
    float dA[3] = {0.0f, 0.0f, 0.0f}, dR = 0.0f, dM = 0.0f;
    const uint32_t slot = (b * 4u + wave) % kGradSlots;
    if (ks.ok) {
        for (uint32_t i = wave; i < ks.n; i += 4u) {  // uniform
            const HotLight L = s_l[i];
            const uint32_t li = hl_uniform(kept.idx[i]);
            const uint64_t lanes = kept.lanes[i];
            adjoint_light(a, L, li, ((lanes >> lane) & 1ull) != 0ull, p, gl, dA, dR, dM, slot);
        }
    } else {
        uint32_t ord2 = 0;
        walk_lights<GRID>(a, s_l, p, valid, cell, [&](const HotLight& L, uint32_t li, bool mine) {
            if ((ord2++ & 3u) != wave) return;
            adjoint_light(a, L, li, mine, p, gl, dA, dR, dM, slot);
        });
    }
    float (*red)[5][64] = as.red[0];
    red[wave][0][lane] = dA[0];
    red[wave][1][lane] = dA[1];
    red[wave][2][lane] = dA[2];
    red[wave][3][lane] = dR;
    red[wave][4][lane] = dM;
    __syncthreads();
    if (valid && lead) {
        float g[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) g[k] = ((red[0][k][lane] + red[1][k][lane]) + red[2][k][lane]) + red[3][k][lane];
        const hl_gptr<float> dp = hl_g(a.d_pixel);
#pragma unroll
        for (int c = 0; c < 3; ++c) dp[c * plane + pix] = g[c] + gl[c] * mixv[c] * p.ao;
        dp[3 * plane + pix] = g[3];
        dp[4 * plane + pix] = g[4];
    }
    __syncthreads();  // red and gl are rewritten by the next item
}

// The same for block tile b with a wave tile per wave (views of at least
// kFusedSplitBlocks block tiles): each wave walks its own tile's lights,
// evaluates all of them in the forward (keeping them in LDS), and the same
// ones in the adjoint.  Only wave-level synchronisation: a wave with no
// geometry returns at once.
template <int FA, int FN, int GRID>
__device__ __forceinline__ void fused_tile_own(const ShadeArgs& a, uint32_t b, ShadeSmem& sm, AdjSmem& as) {
    constexpr int MODE = kModeFusedL2;
    const ViewConst& vc = sm.vc;
    b = hl_uniform(b);
    const uint32_t lane = threadIdx.x & 63u, wave = hl_uniform(threadIdx.x >> 6);
    const uint32_t twl = a.tw_log2, tw = 1u << twl, th = 64u >> twl;
    const uint32_t bx = b % a.block_tiles_x, by = b / a.block_tiles_x;
    const uint32_t x = bx * (4u * tw) + wave * tw + (lane & (tw - 1u));
    const uint32_t y = by * th + (lane >> twl);
    const bool inside = x < a.width && y < a.height;
    const float depth = inside ? hl_load_depth(a.depth, x, y) : 1.0f;
    const bool valid = inside && depth < 1.0f;
    const size_t plane = (size_t)a.width * a.height;
    const size_t pix = (size_t)y * a.width + x;
    float4* out_px = inside ? (float4*)(reinterpret_cast<uint8_t*>(a.out) + (size_t)y * a.out_pitch) + x : nullptr;
    if (inside && !valid) {
        if (a.flags & HL_SHADE_CLEAR_SKY) hl_store_nt(out_px, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        for (int k = 0; k < HL_PIXEL_GRAD_PLANES; ++k) hl_g(a.d_pixel)[k * plane + pix] = 0.0f;
    }
    if (hl_ballot(valid) == 0) return;  // uniform over the wave

    Pixel p;
    uint32_t cell;
    float4 A, E;
    setup_pixel<FA, FN, GRID, MODE>(a, vc, x, y, depth, valid, p, cell, A, E);
    float mixv[3], amb[3];
    ambient_terms(a, p, E, mixv, amb);

    // ---- forward: every light of the tile, kept in LDS for the adjoint
    HotLight* s_l = sm.light[wave];
    KeptLights& kept = as.kept[wave];
    KeptState ks;
    float lo[3] = {0.0f, 0.0f, 0.0f};
    walk_lights<GRID, true>(a, s_l, p, valid, cell, [&](const HotLight& L, uint32_t, bool mine) {
        if (!mine) return;
        Term t;
        if (eval_term<false>(L, p, t)) {
            const float aw = t.atten * t.ndl;
            lo[0] += t.B[0] * (L.p1.x * aw);
            lo[1] += t.B[1] * (L.p1.y * aw);
            lo[2] += t.B[2] * (L.p1.z * aw);
        }
    }, &kept, &ks);
    float o[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = lo[c] + amb[c];

    float gl[3] = {0.0f, 0.0f, 0.0f};
    float l2 = 0.0f;
    if (valid) {
        hl_store_nt(out_px, make_float4(o[0], o[1], o[2], A.w));
        const float4 tg = hl_ld4((const float4*)(reinterpret_cast<const uint8_t*>(a.dout) + (size_t)y * a.dout_pitch) + x);
        const float e0 = o[0] - tg.x, e1 = o[1] - tg.y, e2 = o[2] - tg.z;
        gl[0] = 2.0f * e0;
        gl[1] = 2.0f * e1;
        gl[2] = 2.0f * e2;
        l2 = e0 * e0 + e1 * e1 + e2 * e2;
    }
    const float wl = dpp_sum(l2);
    if (lane == 0) atomicAdd(a.loss, (double)wl);

    // ---- adjoint: the same lights, from LDS when they were kept
    float dA[3] = {0.0f, 0.0f, 0.0f}, dR = 0.0f, dM = 0.0f;
    const uint32_t slot = (b * 4u + wave) % kGradSlots;
    if (ks.ok) {
        for (uint32_t i = 0; i < ks.n; ++i) {  // uniform
            const HotLight L = s_l[i];
            const uint32_t li = hl_uniform(kept.idx[i]);
            const uint64_t lanes = kept.lanes[i];
            adjoint_light(a, L, li, ((lanes >> lane) & 1ull) != 0ull, p, gl, dA, dR, dM, slot);
        }
    } else {
        walk_lights<GRID>(a, s_l, p, valid, cell, [&](const HotLight& L, uint32_t li, bool mine) {
            adjoint_light(a, L, li, mine, p, gl, dA, dR, dM, slot);
        });
    }
    if (valid) {
        const hl_gptr<float> dp = hl_g(a.d_pixel);
#pragma unroll
        for (int c = 0; c < 3; ++c) dp[c * plane + pix] = dA[c] + gl[c] * mixv[c] * p.ao;
        dp[3 * plane + pix] = dR;
        dp[4 * plane + pix] = dM;
    }
}

// Item b of the fused forward + L2 adjoint: a wave tile shared by the block
// on views below kFusedSplitBlocks block tiles, else a block tile.
template <int FA, int FN, int GRID>
__device__ __forceinline__ void fused_tile(const ShadeArgs& a, uint32_t b, ShadeSmem& sm, AdjSmem& as) {
    if (fused_split(a.num_blocks)) fused_tile_split<FA, FN, GRID>(a, b, sm, as);
    else fused_tile_own<FA, FN, GRID>(a, b, sm, as);
}

// ------------------------------------------------------------------ kernels
// Work distribution of the persistent launches: num_items items dealt
// round-robin to kShadeQueues sub-queues (sub-queue q holds items q,
// q + kShadeQueues, ...), each with its own claim counter on its own 128-B
// line, so that the claims of thousands of blocks do not serialise on one
// address.  A block claims from its own sub-queue until it is drained, then
// reads all counters at once (one per thread of the first wave) and moves to
// the next sub-queue with items left; it calls f(item) for every item it
// claims and returns when every sub-queue is drained.  The claim of the next
// item is issued before f runs on the current one, so its round trip overlaps
// the tile's own loads.  Claims past a sub-queue's end only raise its counter.
struct ClaimSmem {
    uint32_t item, q;
};
static_assert(kShadeQueues == 64, "one lane of the first wave per sub-queue");

__device__ __forceinline__ uint32_t queue_len(uint32_t num_items, uint32_t q) {
    return q < num_items ? (num_items - q + kShadeQueues - 1u) / kShadeQueues : 0u;
}

template <class F>
__device__ __forceinline__ void claim_items(uint32_t num_items, uint32_t* __restrict__ queue,
                                            ClaimSmem& cs, F&& f) {
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t q = blockIdx.x % kShadeQueues;
    uint32_t pending = 0;  // thread 0: the claim in flight on sub-queue q
    if (threadIdx.x == 0) pending = atomicAdd(queue + q * kShadeQueueStride, 1u);
    for (;;) {
        if (threadIdx.x == 0) cs.item = pending;
        __syncthreads();
        const uint32_t j = cs.item;
        __syncthreads();
        if (j < queue_len(num_items, q)) {
            if (threadIdx.x == 0) pending = atomicAdd(queue + q * kShadeQueueStride, 1u);
            f(q + j * kShadeQueues);
            continue;
        }
        // drained: the next sub-queue (cyclically after q) with items left
        if (threadIdx.x < 64u) {
            const uint32_t r = threadIdx.x;
            const uint32_t len = queue_len(num_items, r);
            const uint32_t seen =
                len ? __hip_atomic_load(queue + r * kShadeQueueStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                    : 0u;
            const uint64_t left = hl_ballot(seen < len);
            if (threadIdx.x == 0) {
                const uint32_t s0 = (q + 1u) % 64u;  // search from s0 on, wrapping
                const uint64_t rot = s0 == 0u ? left : ((left >> s0) | (left << (64u - s0)));
                cs.q = rot ? (s0 + (uint32_t)__builtin_ctzll(rot)) % 64u : kNone;
            }
        }
        __syncthreads();
        q = cs.q;
        __syncthreads();
        if (q == kNone) return;
        if (threadIdx.x == 0) pending = atomicAdd(queue + q * kShadeQueueStride, 1u);
    }
}

// One view, persistent: the block tiles of the view are the items.  The
// counters of the previous launch on the stream (finished by stream order)
// are zeroed here for the next one, so consecutive launches alternate
// between two counter sets without a separate reset.
// Forward launches on packed G-buffers are held to 96 VGPRs (5 waves per
// SIMD), the shadowed ones and the FP32-G-buffer forward to 128 (4 waves):
// the tile's memory round trips need the waves to hide them.  The pair count
// and the adjoint modes keep their registers (no occupancy request): at 4
// waves they would spill.
template <int MODE, int FA>
constexpr int shade_min_blocks() {
    return MODE == kModeForward && FA != HL_FMT_RGBA32_FLOAT
               ? 5
               : (MODE == kModeForward || MODE == kModeShadowed || MODE == kModeExtended ? 4 : 1);
}

// MODE: a ShadeMode, with kModeOutF16 set for an RGBA16F target
template <int FA, int FN, int GRID, int MODE>
__global__ __launch_bounds__(256, (shade_min_blocks<(MODE & ~kModeOutF16), FA>())) void k_shade(
    ShadeArgs a, uint32_t* __restrict__ queue, uint32_t* __restrict__ queue_next) {
    constexpr int M = MODE & ~kModeOutF16;
    __shared__ ShadeSmem sm;
    __shared__ ClaimSmem cs;
    AdjSmem* as = nullptr;
    if constexpr (M == kModeFusedL2 || M == kModeBackward) {
        __shared__ AdjSmem adj;
        as = &adj;
    }
    if (blockIdx.x == 0 && threadIdx.x < kShadeQueues) queue_next[threadIdx.x * kShadeQueueStride] = 0u;
    load_view_const(a, sm);  // published by the first claim's barrier
    claim_items(shade_items(M, a.num_blocks), queue, cs, [&](uint32_t t) {
        if constexpr (M == kModeFusedL2) fused_tile<FA, FN, GRID>(a, t, sm, *as);
        else shade_tile<FA, FN, GRID, M, (MODE & kModeOutF16) != 0>(a, t, sm, as);
    });
}

// Many views, one persistent launch: the items are the block tiles of the
// views, view after view (heaviest first by the host), described by spans
// (first item, view); an item's view and tile follow from its index.  One
// launch covers the views of one G-buffer format set and one mode (forward,
// or forward + L2 adjoint): the adjoint needs about twice the registers,
// which would halve the forward's occupancy in a shared kernel.  A block
// keeps the item range of its current view; when an item falls outside it
// (rare: each sub-queue's items run through one view at a time) the block
// finds the span by bisection and stages that view's constants in LDS.  The
// view's arguments are read where they lie, through a block-uniform index:
// the loads are scalar, as k_shade's kernel arguments are.
// views is read through the constant address space: the records are not
// written while the launch runs, and their uniform loads become scalar loads.
typedef const __attribute__((address_space(4))) BatchShadeView* ConstViews;

// A launch of one view (nspans == 1) takes that view's arguments from `one`,
// a kernel argument, as k_shade does: read through the view table, the
// arguments the tile uses in its loops were reloaded there (18x the scalar
// loads of k_shade on the same tiles) rather than kept in scalar registers.
template <int FA, int FN, int MODE>
__global__ __launch_bounds__(256, MODE == kModeForward ? 5 : 1) void k_shade_batch(
    ConstViews __restrict__ views, const uint2* __restrict__ spans, uint32_t nspans,
    uint32_t num_items, uint32_t* __restrict__ queue, ShadeArgs one) {
    __shared__ ShadeSmem sm;
    __shared__ ClaimSmem cs;
    AdjSmem* as = nullptr;
    if constexpr (MODE == kModeFusedL2) {
        __shared__ AdjSmem adj;
        as = &adj;
    }
    uint32_t first = 1u, end = 0u, v = 0u;  // item range of the staged view (none yet), its index
    if (nspans == 1u) {
        load_view_const(one, sm);  // published by the first claim's barrier
        claim_items(num_items, queue, cs, [&](uint32_t it) {
            if constexpr (MODE == kModeFusedL2) fused_tile<FA, FN, HL_GRID_CLUSTERS>(one, it, sm, *as);
            else shade_tile<FA, FN, HL_GRID_CLUSTERS, MODE>(one, it, sm, as);
        });
    } else claim_items(num_items, queue, cs, [&](uint32_t it) {
        it = hl_uniform(it);  // block-uniform: every thread read the same claim
        if (it < first || it >= end) {
            uint32_t lo = 0, hi = nspans;  // spans[lo].x <= it < spans[hi].x
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (hl_uniform(spans[mid].x) <= it) lo = mid;
                else hi = mid;
            }
            first = hl_uniform(spans[lo].x);
            end = hl_uniform(spans[lo + 1].x);
            v = hl_uniform(spans[lo].y);
            __syncthreads();  // the previous view's constants are no longer read
            load_view_const(*(const ShadeArgs*)&views[v].sa, sm);
            __syncthreads();
        }
        // v again as a scalar: carried across the claim loop it is not known
        // to be uniform
        const ShadeArgs& ra = *(const ShadeArgs*)&views[hl_uniform(v)].sa;
        if constexpr (MODE == kModeFusedL2) fused_tile<FA, FN, HL_GRID_CLUSTERS>(ra, it - first, sm, *as);
        else shade_tile<FA, FN, HL_GRID_CLUSTERS, MODE>(ra, it - first, sm, as);
    });
    // The last block out clears the claim counters for the next launch, so a
    // captured graph replays this launch unchanged.  Blocks that have finished
    // claiming are counted in two levels so that no single address takes an
    // atomic from every block (those serialise): word 1 of sub-queue g's line
    // counts the blocks with blockIdx % kShadeQueues == g, and the last of
    // them adds to word 2 of line 0, which counts the finished groups.  No
    // fence: each claim's atomic has returned before these are issued.
    __shared__ uint32_t last;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = blockIdx.x % kShadeQueues;
        const uint32_t group_blocks = (gridDim.x - g + kShadeQueues - 1u) / kShadeQueues;
        const uint32_t groups = gridDim.x < kShadeQueues ? gridDim.x : kShadeQueues;
        uint32_t l = 0u;
        if (atomicAdd(queue + g * kShadeQueueStride + 1u, 1u) == group_blocks - 1u)
            l = atomicAdd(queue + 2u, 1u) == groups - 1u ? 1u : 0u;
        last = l;
    }
    __syncthreads();
    if (last && threadIdx.x < kShadeQueues) {
        queue[threadIdx.x * kShadeQueueStride] = 0u;
        queue[threadIdx.x * kShadeQueueStride + 1u] = 0u;
        if (threadIdx.x == 0) queue[2] = 0u;
    }
}

template <int FA, int FN>
static void launch_batch_fmt(hipStream_t s, int mode, const BatchShadeView* views, const uint2* spans,
                             uint32_t nspans, uint32_t num_items, uint32_t* queue, uint32_t grid_blocks,
                             const ShadeArgs& one) {
    if (mode == kModeFusedL2)
        k_shade_batch<FA, FN, kModeFusedL2><<<grid_blocks, 256, 0, s>>>((ConstViews)views, spans, nspans, num_items,
                                                                        queue, one);
    else
        k_shade_batch<FA, FN, kModeForward><<<grid_blocks, 256, 0, s>>>((ConstViews)views, spans, nspans, num_items,
                                                                        queue, one);
}

template <int FA, int FN>
static uint32_t batch_blocks_per_cu(int mode) {
    int n = 0;
    const hipError_t e = mode == kModeFusedL2
                             ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_shade_batch<FA, FN, kModeFusedL2>, 256, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_shade_batch<FA, FN, kModeForward>, 256, 0);
    return e == hipSuccess && n > 0 ? (uint32_t)n : 1u;
}

uint32_t shade_batch_blocks_per_cu(int gbuf_set, int mode) {
    switch (gbuf_set) {
        case kSetRef: return batch_blocks_per_cu<HL_FMT_RGBA8_UNORM, HL_FMT_RGBA16_FLOAT>(mode);
        case kSet16: return batch_blocks_per_cu<HL_FMT_RGBA16_FLOAT, HL_FMT_RGBA16_FLOAT>(mode);
        default: return batch_blocks_per_cu<HL_FMT_RGBA32_FLOAT, HL_FMT_RGBA32_FLOAT>(mode);
    }
}

hipError_t launch_shade_batch(hipStream_t s, int gbuf_set, int mode, const BatchShadeView* views,
                              const uint2* spans, uint32_t nspans, uint32_t num_items, uint32_t* queue,
                              uint32_t grid_blocks, const ShadeArgs* one) {
    if (num_items == 0) return hipSuccess;
    if ((nspans == 1u) != (one != nullptr)) return hipErrorInvalidValue;
    const ShadeArgs a = one ? *one : ShadeArgs{};
    switch (gbuf_set) {
        case kSetRef:
            launch_batch_fmt<HL_FMT_RGBA8_UNORM, HL_FMT_RGBA16_FLOAT>(s, mode, views, spans, nspans, num_items,
                                                                    queue, grid_blocks, a);
            break;
        case kSet16:
            launch_batch_fmt<HL_FMT_RGBA16_FLOAT, HL_FMT_RGBA16_FLOAT>(s, mode, views, spans, nspans, num_items,
                                                                     queue, grid_blocks, a);
            break;
        default:
            launch_batch_fmt<HL_FMT_RGBA32_FLOAT, HL_FMT_RGBA32_FLOAT>(s, mode, views, spans, nspans, num_items,
                                                                     queue, grid_blocks, a);
            break;
    }
    return hipGetLastError();
}

// Resident blocks of a persistent launch: what the kernel's occupancy allows
// on every CU (queried once per kernel), at most one per item.
template <int FA, int FN, int GRID, int MODE>
static uint32_t persistent_blocks(uint32_t items) {
    static int per_cu = -1, cus = 0;
    if (per_cu < 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_shade<FA, FN, GRID, MODE>, 256, 0) !=
                hipSuccess || n < 1)
            n = 4, cus = cus > 0 ? cus : 256;
        per_cu = n;
    }
    const uint32_t b = (uint32_t)(per_cu * cus);
    return items < b ? items : b;
}

template <int FA, int FN, int MODE>
static hipError_t launch_grid(hipStream_t s, const ShadeArgs& a, uint32_t* queue, uint32_t* queue_next) {
    if (a.grid_mode == HL_GRID_TILES)
        k_shade<FA, FN, HL_GRID_TILES, MODE>
            <<<persistent_blocks<FA, FN, HL_GRID_TILES, MODE>(shade_items(MODE & ~kModeOutF16, a.num_blocks)), 256,
               0, s>>>(
                a, queue, queue_next);
    else
        k_shade<FA, FN, HL_GRID_CLUSTERS, MODE>
            <<<persistent_blocks<FA, FN, HL_GRID_CLUSTERS, MODE>(shade_items(MODE & ~kModeOutF16, a.num_blocks)),
               256, 0, s>>>(
                a, queue, queue_next);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_fmt(hipStream_t s, int set, const ShadeArgs& a, uint32_t* q, uint32_t* qn) {
    switch (set) {
        case kSetRef: return launch_grid<HL_FMT_RGBA8_UNORM, HL_FMT_RGBA16_FLOAT, MODE>(s, a, q, qn);
        case kSet16: return launch_grid<HL_FMT_RGBA16_FLOAT, HL_FMT_RGBA16_FLOAT, MODE>(s, a, q, qn);
        default: return launch_grid<HL_FMT_RGBA32_FLOAT, HL_FMT_RGBA32_FLOAT, MODE>(s, a, q, qn);
    }
}

hipError_t launch_shade(hipStream_t s, int gbuf_set, int mode, const ShadeArgs& a, uint32_t* queue,
                        uint32_t* queue_next) {
    if (a.num_blocks == 0) return hipSuccess;
    if (a.out_f16) {  // the modes that write the radiance alone take an RGBA16F target
        switch (mode) {
            case kModeForward: return launch_fmt<kModeForward | kModeOutF16>(s, gbuf_set, a, queue, queue_next);
            case kModeShadowed: return launch_fmt<kModeShadowed | kModeOutF16>(s, gbuf_set, a, queue, queue_next);
            case kModeExtended: return launch_fmt<kModeExtended | kModeOutF16>(s, gbuf_set, a, queue, queue_next);
            default: return hipErrorInvalidValue;
        }
    }
    switch (mode) {
        case kModeForward: return launch_fmt<kModeForward>(s, gbuf_set, a, queue, queue_next);
        case kModeCount: return launch_fmt<kModeCount>(s, gbuf_set, a, queue, queue_next);
        case kModeBackward: return launch_fmt<kModeBackward>(s, gbuf_set, a, queue, queue_next);
        case kModeFusedL2: return launch_fmt<kModeFusedL2>(s, gbuf_set, a, queue, queue_next);
        case kModeShadowed: return launch_fmt<kModeShadowed>(s, gbuf_set, a, queue, queue_next);
        case kModeExtended: return launch_fmt<kModeExtended>(s, gbuf_set, a, queue, queue_next);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------ light gradient I/O
// FP64 accumulators (kGradSlots copies of 4 per light of the launch, slot
// major) -> their sums as FP32 outputs indexed by the light's pool index
// (subset[i], or i).
__global__ void k_light_grads_to_f32(const double* __restrict__ acc, uint32_t n,
                                     const uint32_t* __restrict__ subset, float* __restrict__ d_pos,
                                     float* __restrict__ d_int) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double g[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (uint32_t s = 0; s < kGradSlots; ++s)
#pragma unroll
        for (int c = 0; c < 4; ++c) g[c] += acc[4 * ((size_t)s * n + i) + c];
    const uint32_t o = subset ? subset[i] : i;
    d_pos[3 * (size_t)o + 0] = (float)g[0];
    d_pos[3 * (size_t)o + 1] = (float)g[1];
    d_pos[3 * (size_t)o + 2] = (float)g[2];
    d_int[o] = (float)g[3];
}

// Every adjoint view of a plan at once: view blockIdx.y, its words striding
// over the x blocks.
__global__ void k_grad_views_zero(const GradView* __restrict__ views) {
    const GradView v = views[blockIdx.y];
    const size_t acc_words = 2 * grad_accum_doubles(v.n), pos_words = 3 * (size_t)v.pool;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < acc_words || i < pos_words; i += stride) {
        if (i < acc_words) reinterpret_cast<uint32_t*>(v.acc)[i] = 0u;
        if (i < pos_words) v.d_pos[i] = 0.0f;
        if (i < v.pool) v.d_int[i] = 0.0f;
        if (i == 0 && v.loss) *v.loss = 0.0;
    }
}

__global__ void k_grad_views_to_f32(const GradView* __restrict__ views) {
    const GradView v = views[blockIdx.y];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.n) return;
    double g[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (uint32_t s = 0; s < kGradSlots; ++s)
#pragma unroll
        for (int c = 0; c < 4; ++c) g[c] += v.acc[4 * ((size_t)s * v.n + i) + c];
    const uint32_t o = v.subset ? v.subset[i] : i;
    v.d_pos[3 * (size_t)o + 0] = (float)g[0];
    v.d_pos[3 * (size_t)o + 1] = (float)g[1];
    v.d_pos[3 * (size_t)o + 2] = (float)g[2];
    v.d_int[o] = (float)g[3];
}

// The Vogel tables of kVogelFloats floats: block n, thread i.
__global__ void k_vogel_table(float* __restrict__ t) {
#pragma clang fp contract(off)
    const uint32_t n = blockIdx.x, i = threadIdx.x;
    if (n == 0u) {
        float sa, ca;
        sh_sincos((float)i * 2.39996323f, sa, ca);
        t[i] = ca;
        t[kVogelMax + i] = sa;
    }
    if (i < n) t[kVogelRadii + (size_t)n * kVogelMax + i] = sqrtf((float)i + 0.5f) / sqrtf((float)n);
}

// sample_rc_irradiance per non-sky pixel ahead of the extended shading launch,
// from the N and world position that kernel computes (exact normalize, the
// same row sums over inv_view_proj): the atlas walk stays out of the light
// loop's register budget.  One thread per pixel, 64 x 4 pixels per block.
template <int FN>
__global__ __launch_bounds__(256) void k_rc_irradiance(ShadeArgs a) {
#pragma clang fp contract(off)
    const uint32_t x = blockIdx.x * 64u + (threadIdx.x & 63u), y = blockIdx.y * 4u + (threadIdx.x >> 6);
    if (x >= a.width || y >= a.height) return;
    const float depth = hl_load_depth(a.depth, x, y);
    if (!(depth < 1.0f)) return;
    const float4 Nr = hl_load_rgba<FN>(a.normal, x, y);
    const float3 n = hl_x_normalize(Nr.x, Nr.y, Nr.z);
    const float u = ((float)x + 0.5f) / (float)a.width;
    const float v = ((float)y + 0.5f) / (float)a.height;
    const float ndx = u * 2.0f - 1.0f, ndy = 1.0f - v * 2.0f;
    const float* m = a.cam->inv_view_proj;
    const float hw = hl_x_mat_row(m, 3, ndx, ndy, depth, 1.0f);
    const float wx = hl_x_mat_row(m, 0, ndx, ndy, depth, 1.0f) / hw;
    const float wy = hl_x_mat_row(m, 1, ndx, ndy, depth, 1.0f) / hw;
    const float wz = hl_x_mat_row(m, 2, ndx, ndy, depth, 1.0f) / hw;
    const float3 rc = rc_irradiance(a, wx, wy, wz, n.x, n.y, n.z);
    hl_store_nt(const_cast<float4*>(a.rc_irr) + (size_t)y * a.width + x, make_float4(rc.x, rc.y, rc.z, 0.0f));
}

hipError_t launch_rc_irradiance(hipStream_t s, int gbuf_set, const ShadeArgs& a) {
    const dim3 grid((a.width + 63u) / 64u, (a.height + 3u) / 4u);
    if (gbuf_set == kSet32) k_rc_irradiance<HL_FMT_RGBA32_FLOAT><<<grid, 256, 0, s>>>(a);
    else k_rc_irradiance<HL_FMT_RGBA16_FLOAT><<<grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_vogel_table(hipStream_t s, float* t) {
    k_vogel_table<<<kVogelMax + 1, kVogelMax, 0, s>>>(t);
    return hipGetLastError();
}

hipError_t launch_light_grads_to_f32(hipStream_t s, const double* acc, uint32_t n,
                                     const uint32_t* subset, float* d_pos, float* d_int) {
    if (n == 0) return hipSuccess;
    k_light_grads_to_f32<<<(n + 255) / 256, 256, 0, s>>>(acc, n, subset, d_pos, d_int);
    return hipGetLastError();
}

hipError_t launch_grad_views_zero(hipStream_t s, const GradView* views, uint32_t nv, size_t max_words) {
    if (nv == 0) return hipSuccess;
    size_t blocks = (max_words + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    k_grad_views_zero<<<dim3((unsigned)blocks, nv), 256, 0, s>>>(views);
    return hipGetLastError();
}

hipError_t launch_grad_views_to_f32(hipStream_t s, const GradView* views, uint32_t nv, uint32_t max_n) {
    if (nv == 0 || max_n == 0) return hipSuccess;
    k_grad_views_to_f32<<<dim3((max_n + 255) / 256, nv), 256, 0, s>>>(views);
    return hipGetLastError();
}

}  // namespace hl
