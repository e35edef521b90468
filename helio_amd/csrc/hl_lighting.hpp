// hl_lighting.hpp -- the per-(pixel, light) shading shared by the deferred
// pass (hl_shade.hip) and the forward-lit fragments (hl_forward.hip): the
// pixel and term records, the range test, pbr_direct_light with fast
// arithmetic and exact decisions, and the wave-cooperative walk over a cell's
// light list (lights refined against the wave's AABB and staged in LDS).
#pragma once

#include "hl_internal.hpp"

namespace hl {

constexpr int kStage = 64;  // LDS light records per wave and chunk
constexpr float kPi = 3.14159265359f;
constexpr float kInvPi = 1.0f / 3.14159265359f;
constexpr float kThresh = 0.002f;  // deferred_lighting.wgsl:781

struct Pixel {
    float wx, wy, wz;
    float nx, ny, nz;
    float vx, vy, vz;
    float f0[3];
    float alb[3];
    float rough, metal, ao;
    float ndv, kk, kk1, gv, ev, one_m;  // kk1 = kk + 0.0001
    float kd[3];     // (1 - metallic) albedo / pi
    float a2, a2gv;  // roughness^4 and roughness^4 G1(NdV)
    bool f0_free[3];
};

// Per-(pixel, light) quantities of pbr_direct_light (deferred_lighting.wgsl:699-820).
struct Term {
    float lx, ly, lz;
    float rad[3];
    float ndl;
    float d, inv_d, att0, fall, s, dsdcos, atten;
    float hx, hy, hz, inv_h;
    float hdv, u, P, ndh, a2, den, Q, D, el, gl, G, den2;
    float F[3], B[3];
};

// Range test of deferred_lighting.wgsl:1128-1131 (dist > range rejects)
// from the fast squared distance; inside the +-1e-5 relative band around
// range^2, which covers the rounding of both sides many times over, the exact
// restatement decides.  The band is rare: the exact branch is taken only when
// some lane of the wave needs it.
__device__ __forceinline__ bool range_test(const HotLight& L, float tx, float ty, float tz, float d2) {
    const float r2 = L.p0.w * L.p0.w;
    bool in = d2 < r2 * (1.0f - 1e-5f);
    const bool band = !in && !(d2 > r2 * (1.0f + 1e-5f));
    if (hl_ballot(band)) {
        hl_rare();
        if (band) in = !(sqrtf(hl_x_dot3(tx, ty, tz, tx, ty, tz)) > L.p0.w);
    }
    return in;
}

// The geometry of a term near the GGX peak in the shader's order
// (deferred_lighting.wgsl:699-760, pbr_eval.wgsl:12-18): L from the light,
// H = normalize(V + L) with the pixel's exact V, the three dot products and
// the GGX denominator, IEEE throughout.  False when N.L is exactly zero there
// (the shader returns before the BRDF).
__device__ __forceinline__ bool exact_peak(const HotLight& L, const Pixel& p, Term& t) {
#pragma clang fp contract(off)
    const float3 l = __builtin_isinf(L.p0.w)
                         ? hl_x_normalize_r(-L.p2.x, -L.p2.y, -L.p2.z)
                         : hl_x_normalize_r(L.p0.x - p.wx, L.p0.y - p.wy, L.p0.z - p.wz);
    const float ndl = fmaxf(hl_x_dot3(p.nx, p.ny, p.nz, l.x, l.y, l.z), 0.0f);
    if (ndl == 0.0f) return false;
    float hl;
    const float3 h = hl_x_normalize_r(p.vx + l.x, p.vy + l.y, p.vz + l.z, &hl);
    t.lx = l.x;
    t.ly = l.y;
    t.lz = l.z;
    t.ndl = ndl;
    t.inv_h = hl_f_rcp(hl);
    t.hx = h.x;
    t.hy = h.y;
    t.hz = h.z;
    t.hdv = fmaxf(hl_x_dot3(h.x, h.y, h.z, p.vx, p.vy, p.vz), 0.0f);
    t.ndh = fmaxf(hl_x_dot3(p.nx, p.ny, p.nz, h.x, h.y, h.z), 0.0f);
    t.den = t.ndh * t.ndh * (p.a2 - 1.0f) + 1.0f;
    return true;
}

// Fast arithmetic, exact decisions.  Returns false when the term is exactly
// zero in the restated shader.
// kCut: apply the all(radiance < 0.002) early-out of the deferred shader
// (deferred_lighting.wgsl:781); the forward-lit shader has none
// (forward_lit.wgsl:374-438).
// Near the GGX peak of a smooth surface the denominator NdH^2 (a2 - 1) + 1
// is a difference of terms near 1 and as small as a2 (6e-6 for roughness
// 0.05): a last-bit error of the fast L or H (a few 1e-7 in NdH, N and V being
// exact from the pixel set-up) moves D by up to ~1e-6 / den relative.  Below
// kDenExact L, H, N.L, H.V, N.H and the denominator are recomputed in the
// shader's order with IEEE operations.
constexpr float kDenExact = 0.02f;

template <bool kFull, bool kCut = true>
__device__ __forceinline__ bool eval_term(const HotLight& L, const Pixel& p, Term& t) {
    // Both branches compute locals and the record is written once below:
    // fields stored in different branches were merged into one store through
    // a selected address, which kept the record on the stack.
    float lx, ly, lz, d = 0.0f, inv_d = 0.0f, att0 = 1.0f, fall = 1.0f, s = 1.0f, dsdcos = 0.0f;
    if (__builtin_isinf(L.p0.w)) {  // directional
        const float dx = -L.p2.x, dy = -L.p2.y, dz = -L.p2.z;
        const float inv = hl_f_rsq(dx * dx + dy * dy + dz * dz);
        lx = dx * inv;
        ly = dy * inv;
        lz = dz * inv;
        if (kCut && L.p4.z < kThresh) return false;
    } else {
        const float tx = L.p0.x - p.wx, ty = L.p0.y - p.wy, tz = L.p0.z - p.wz;
        const float d2 = __builtin_fmaf(tx, tx, __builtin_fmaf(ty, ty, tz * tz));
        // The first decision in one band test on d2 (p4.xy, built by
        // make_hot_light): below p4.x the pixel is in range and, with kCut,
        // the unshadowed radiance is at least 0.002 (deferred_lighting.wgsl:
        // 1128-1131 and 781); above p4.y the term is zero; in between the
        // exact restatement decides.
        bool in = d2 < L.p4.x;
        const bool band = !in && !(d2 > L.p4.y);
        if (hl_ballot(band)) {
            hl_rare();
            if (band) {
                bool in_range;
                const float m = hl_x_radiance_max(L, p.wx, p.wy, p.wz, &in_range);
                in = in_range && (!kCut || !(m < kThresh));
            }
        }
        if (!in) return false;
        inv_d = hl_f_rsq(d2);
        d = d2 * inv_d;
        lx = tx * inv_d;
        ly = ty * inv_d;
        lz = tz * inv_d;
        att0 = hl_f_rcp(d2 + 0.0001f);
        const float nd2 = d2 * L.p3.x;  // (dist / range)^2
        fall = fmaxf(0.0f, 1.0f - nd2 * nd2);
        if (L.p2.w != kNoCone) {
            // smoothstep(cos outer, cos inner, cos_a) (deferred_lighting.wgsl:
            // 739-742) divides by cos(inner) - cos(outer) (0.005 for a 10-degree
            // penumbra), which would amplify a last-bit error of a fast L ~200x.
            // Outside the penumbra by more than the margin the factor is exactly
            // 0 (the term vanishes) or 1; inside, it is evaluated exactly, and
            // so is the 0.002 cut, which the factor lowers.
            const float cf = -(lx * L.p2.x + ly * L.p2.y + lz * L.p2.z);
            const bool ordered = L.p2.w < L.p1.w;
            if (ordered && cf < L.p2.w - L.p3.z) return false;
            if (!(ordered && cf > L.p1.w + L.p3.z)) {
                hl_rare();
                float cos_a;
                {
#pragma clang fp contract(off)
                    const float dist = sqrtf(hl_x_dot3(tx, ty, tz, tx, ty, tz));
                    const float ex = tx / dist, ey = ty / dist, ez = tz / dist;
                    cos_a = hl_x_dot3(-ex, -ey, -ez, L.p2.x, L.p2.y, L.p2.z);
                }
                s = hl_x_smoothstep(L.p2.w, L.p1.w, cos_a);
                if (kCut) {
                    bool in_range;
                    if (hl_x_radiance_max(L, p.wx, p.wy, p.wz, &in_range) < kThresh) return false;
                }
                if (kFull) {
                    const float inv_w = hl_f_rcp(L.p1.w - L.p2.w);
                    const float tt = (cos_a - L.p2.w) * inv_w;
                    dsdcos = (tt > 0.0f && tt < 1.0f) ? 6.0f * tt * (1.0f - tt) * inv_w : 0.0f;
                }
            }
        }
    }
    // directional: att0 = fall = s = 1, so atten = 1 and rad = p1 exactly
    const float atten = att0 * fall * s;
    t.lx = lx;
    t.ly = ly;
    t.lz = lz;
    t.d = d;
    t.inv_d = inv_d;
    t.att0 = att0;
    t.fall = fall;
    t.s = s;
    t.dsdcos = dsdcos;
    t.atten = atten;
    if (kFull) {
        t.rad[0] = L.p1.x * atten;
        t.rad[1] = L.p1.y * atten;
        t.rad[2] = L.p1.z * atten;
    }
    t.ndl = fmaxf(p.nx * t.lx + p.ny * t.ly + p.nz * t.lz, 0.0f);
    if (t.ndl == 0.0f) return false;
    // H = normalize(V + L) and N.H in the shader's order: for smooth surfaces
    // (roughness^4 ~ 1e-5) the GGX denominator near the peak is as small as
    // one ulp of N.H, so a rearranged N.H (e.g. (N.V + N.L) / |V + L|) moves D
    // by percents there.
    const float hx = p.vx + t.lx, hy = p.vy + t.ly, hz = p.vz + t.lz;
    const float inv_h = hl_f_rsq(hx * hx + hy * hy + hz * hz);
    t.inv_h = inv_h;
    t.hx = hx * inv_h;
    t.hy = hy * inv_h;
    t.hz = hz * inv_h;
    t.hdv = fmaxf(t.hx * p.vx + t.hy * p.vy + t.hz * p.vz, 0.0f);
    t.ndh = fmaxf(p.nx * t.hx + p.ny * t.hy + p.nz * t.hz, 0.0f);
    t.den = t.ndh * t.ndh * (p.a2 - 1.0f) + 1.0f;
    if (hl_ballot(t.den < kDenExact)) {
        hl_rare();
        if (t.den < kDenExact && !exact_peak(L, p, t)) return false;
    }
    t.u = fminf(fmaxf(1.0f - t.hdv, 0.0f), 1.0f);
    const float u2 = t.u * t.u;
    t.P = u2 * u2 * t.u;
    t.a2 = p.a2;
    t.Q = kPi * t.den * t.den + 0.0001f;
    t.el = t.ndl * (1.0f - p.kk) + p.kk1;
    t.den2 = 4.0f * p.ndv * t.ndl + 0.0001f;
    float spec;
    if (kFull) {
        t.D = p.a2 * hl_f_rcp(t.Q);
        t.gl = t.ndl * hl_f_rcp(t.el);
        t.G = p.gv * t.gl;
        spec = t.D * t.G * hl_f_rcp(t.den2);
    } else {
        // D G / den2 with one reciprocal: a2 gv NdL / (Q el den2); for
        // roughness in [0, 1] each factor lies in [1e-4, 4.1]
        spec = p.a2gv * t.ndl * hl_f_rcp(t.Q * t.el * t.den2);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        t.F[c] = p.f0[c] + (1.0f - p.f0[c]) * t.P;
        t.B[c] = p.kd[c] + t.F[c] * (spec - p.kd[c]);  // (1 - F) kd + spec F
    }
    return true;
}

// The lights a wave's walk visited, kept for a second pass over the same
// lights (the fused adjoint): their records stay where the walk staged them,
// compacted in visiting order (s_l[0, n)), with each one's index and the lanes
// of its cell group.  ok turns false when they do not fit in kStage slots.
struct KeptLights {
    uint32_t idx[kStage];
    uint64_t lanes[kStage];
};
struct KeptState {
    uint32_t n = 0;
    bool ok = true;
};

// Walks the lights of every distinct cell among the wave's valid pixels:
// per cell, refines its list against the AABB of those pixels and calls
// f(light, index, mine) for each light that passes, in list order, uniformly
// across the wave.  The cell's list is streamed 64 entries at a time, one per
// lane; a lane whose light passes stores the record in an LDS slot (its own,
// or with kept the next free one of the kept range), and the wave then visits
// the set bits of the ballot (ascending lanes, i.e. list order) with uniform
// reads of those slots.  a provides the grid (lists, counts, offsets,
// lists_capacity) and the light records (lights, num_lights).
template <int GRID, bool kKeep = false, class A, class F>
__device__ __forceinline__ void walk_lights(const A& a, HotLight* s_l, const Pixel& p, bool valid,
                                            uint32_t cell, F&& f, KeptLights* kept = nullptr,
                                            KeptState* ks = nullptr) {
    const uint32_t lane = hl_lane();
    uint64_t todo = hl_ballot(valid);
    while (todo) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(todo);
        const uint32_t ucell = hl_readlane(cell, leader);
        const bool mine = valid && cell == ucell;
        const uint64_t group = hl_ballot(mine);
        todo &= ~group;
        const float bx0 = dpp_min_if(mine, p.wx), bx1 = dpp_max_if(mine, p.wx);
        const float by0 = dpp_min_if(mine, p.wy), by1 = dpp_max_if(mine, p.wy);
        const float bz0 = dpp_min_if(mine, p.wz), bz1 = dpp_max_if(mine, p.wz);
        uint32_t cnt = hl_uniform(hl_g(a.counts)[ucell]);
        uint32_t base = GRID == HL_GRID_TILES ? ucell * HL_MAX_LIGHTS_PER_TILE
                                              : hl_uniform(hl_g(a.offsets)[ucell]);
        if (base >= a.lists_capacity) cnt = 0;
        else if (cnt > a.lists_capacity - base) cnt = a.lists_capacity - base;
        const uint32_t* list = a.lists + base;
        for (uint32_t j0 = 0; j0 < cnt; j0 += 64u) {
            const uint32_t j = j0 + lane;
            bool pass = false;
            uint32_t idx = 0;
            [[maybe_unused]] float4 q, l1, l2, l3, l4;  // kKeep: staged after the ballot
            if (j < cnt) {
                idx = hl_g(list)[j];
                if (idx < a.num_lights) {
                    // the whole record at once, so that staging it does not
                    // wait on a second load after the test
                    const HotLight* hp = a.lights + idx;
                    const float4 r0 = hl_ld4(&hp->p0), r1 = hl_ld4(&hp->p1), r2 = hl_ld4(&hp->p2),
                                 r3 = hl_ld4(&hp->p3), r4 = hl_ld4(&hp->p4);
                    const float dx = fmaxf(fmaxf(bx0 - r0.x, r0.x - bx1), 0.0f);
                    const float dy = fmaxf(fmaxf(by0 - r0.y, r0.y - by1), 0.0f);
                    const float dz = fmaxf(fmaxf(bz0 - r0.z, r0.z - bz1), 0.0f);
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    pass = d2 <= r0.w * r0.w * 1.0001f + 1e-12f;
                    if constexpr (kKeep) {
                        q = r0;
                        l1 = r1;
                        l2 = r2;
                        l3 = r3;
                        l4 = r4;
                    } else if (pass) {  // the lane's own slot, stored while the rest are tested
                        s_l[lane].p0 = r0;
                        s_l[lane].p1 = r1;
                        s_l[lane].p2 = r2;
                        s_l[lane].p3 = r3;
                        s_l[lane].p4 = r4;
                    }
                }
            }
            uint64_t m = hl_ballot(pass);
            // kKeep: the kept range's next slots in lane order while they fit,
            // else the lane's own
            bool keep = false;
            uint32_t first = 0;
            if constexpr (kKeep) {
                const uint32_t np = (uint32_t)__popcll(m);
                if (ks->ok && ks->n + np > (uint32_t)kStage) ks->ok = false;
                keep = ks->ok;
                first = ks->n;
                if (keep) ks->n += np;
                if (pass) {
                    const uint32_t slot = keep ? first + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)) : lane;
                    s_l[slot].p0 = q;
                    s_l[slot].p1 = l1;
                    s_l[slot].p2 = l2;
                    s_l[slot].p3 = l3;
                    s_l[slot].p4 = l4;
                    if (keep) {
                        kept->idx[slot] = idx;
                        kept->lanes[slot] = group;
                    }
                }
            }
            hl_wave_sync();
            uint32_t slot = first;
            while (m) {
                const uint32_t k = (uint32_t)__builtin_ctzll(m);
                m &= m - 1u;
                const HotLight L = s_l[keep ? slot++ : k];
                f(L, hl_readlane(idx, k), mine);
            }
            hl_wave_sync();
        }
    }
}

}  // namespace hl
