// hl_shade.hip -- deferred shading (fs_main subset) and its adjoint for gfx950.
//
// Launch shape: 256-thread blocks covering 32x8 pixels; each wavefront owns an
// 8x8 pixel tile.  A wavefront
//   1. loads depth and the G-buffer (coalesced rows of 32 px per block),
//      reconstructs the world position exactly and finds each pixel's cell;
//   2. for each distinct cell among its pixels: reduces the AABB of those
//      pixels' world positions, streams the cell's list in chunks of 64, tests
//      each light sphere against the AABB, and compacts the survivors into an
//      LDS list with ballot + prefix popcount (order preserved);
//   3. walks the LDS list with a wave-uniform index (light records come in
//      through scalar loads) and evaluates pbr_direct_light for its pixels.
// Every pixel therefore sums exactly the lights of its own cell list that pass
// the range test, in list order, as deferred_lighting.wgsl:1125-1145 does.
#include "hl_internal.hpp"

namespace hl {

constexpr int kLdsCap = 256;
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
    float ndv, kk, gv, one_m;
};

// pbr_direct_light for one light, fast arithmetic with exact decisions.
// Returns false when the term is exactly zero in the restated shader; *in_range
// reports the range test of deferred_lighting.wgsl:1128-1131.
struct Term {
    float lx, ly, lz;    // L
    float rad[3];        // radiance
    float ndl;
    float d, inv_d, att0, fall, s, dsdcos;
    float hx, hy, hz, inv_h;  // H (normalized) and 1/|V+L|
    float hdv, u, P, ndh, a2, den, Q, D, el, gl, G, den2;
    float F[3], B[3];
};

template <bool kFull>
__device__ __forceinline__ bool eval_term(const HotLight& L, const Pixel& p, Term& t,
                                          bool& in_range) {
    const bool dir = __builtin_isinf(L.p0.w);
    if (dir) {
        in_range = true;
        float dx = -L.p2.x, dy = -L.p2.y, dz = -L.p2.z;
        float inv = hl_f_rsq(dx * dx + dy * dy + dz * dz);
        t.lx = dx * inv;
        t.ly = dy * inv;
        t.lz = dz * inv;
        t.rad[0] = L.p1.x;
        t.rad[1] = L.p1.y;
        t.rad[2] = L.p1.z;
        t.att0 = 1.0f;
        t.fall = 1.0f;
        t.s = 1.0f;
        t.dsdcos = 0.0f;
        t.d = 0.0f;
        t.inv_d = 0.0f;
    } else {
        const float tx = L.p0.x - p.wx, ty = L.p0.y - p.wy, tz = L.p0.z - p.wz;
        const float d2 = tx * tx + ty * ty + tz * tz;
        const float r = L.p0.w;
        const float r2 = r * r;
        bool in;
        if (d2 < r2 * (1.0f - 1e-5f)) in = true;
        else if (d2 > r2 * (1.0f + 1e-5f)) in = false;
        else in = !(sqrtf(hl_x_dot3(tx, ty, tz, tx, ty, tz)) > r);
        in_range = in;
        if (!in) return false;
        const float inv_d = hl_f_rsq(d2);
        const float d = d2 * inv_d;
        t.d = d;
        t.inv_d = inv_d;
        t.lx = tx * inv_d;
        t.ly = ty * inv_d;
        t.lz = tz * inv_d;
        t.att0 = hl_f_rcp(d2 + 0.0001f);
        const float nd = d * hl_f_rcp(r);
        const float nd2 = nd * nd;
        t.fall = fmaxf(0.0f, 1.0f - nd2 * nd2);
        t.s = 1.0f;
        t.dsdcos = 0.0f;
        if (L.p2.w != kNoCone) {
            // The cone factor divides by cos(inner) - cos(outer) (0.005 for a
            // 10-degree cone), which would amplify the last-bit error of a fast
            // L ~200x; it is evaluated exactly as deferred_lighting.wgsl:739-742.
            float cos_a;
            {
#pragma clang fp contract(off)
                const float dist = sqrtf(hl_x_dot3(tx, ty, tz, tx, ty, tz));
                const float ex = tx / dist, ey = ty / dist, ez = tz / dist;
                cos_a = hl_x_dot3(-ex, -ey, -ez, L.p2.x, L.p2.y, L.p2.z);
            }
            t.s = hl_x_smoothstep(L.p2.w, L.p1.w, cos_a);
            if (kFull) {
                const float inv_w = hl_f_rcp(L.p1.w - L.p2.w);
                const float tt = (cos_a - L.p2.w) * inv_w;
                t.dsdcos = (tt > 0.0f && tt < 1.0f) ? 6.0f * tt * (1.0f - tt) * inv_w : 0.0f;
            }
        }
        const float atten = t.att0 * t.fall * t.s;
        t.rad[0] = L.p1.x * atten;
        t.rad[1] = L.p1.y * atten;
        t.rad[2] = L.p1.z * atten;
        const float m = fmaxf(fmaxf(t.rad[0], t.rad[1]), t.rad[2]);
        // error bound of the fast radiance (rcp/rsq of att0, |fall| near 0)
        const float ci = fmaxf(fmaxf(L.p1.x, L.p1.y), L.p1.z);
        const float err = 8.0f * ci * t.att0 * 4e-6f * t.s + m * 4e-6f + 1e-12f;
        if (m < kThresh - err) return false;
        if (m < kThresh + err) {  // ambiguous: decide with the exact restatement
            bool dummy;
            if (hl_x_radiance_max(L, p.wx, p.wy, p.wz, &dummy) < kThresh) return false;
        }
    }
    const float ndl_raw = p.nx * t.lx + p.ny * t.ly + p.nz * t.lz;
    t.ndl = fmaxf(ndl_raw, 0.0f);
    if (t.ndl == 0.0f) return false;
    if (dir && fmaxf(fmaxf(t.rad[0], t.rad[1]), t.rad[2]) < kThresh) return false;
    // H = normalize(V + L)
    const float hx = p.vx + t.lx, hy = p.vy + t.ly, hz = p.vz + t.lz;
    const float inv_h = hl_f_rsq(hx * hx + hy * hy + hz * hz);
    t.inv_h = inv_h;
    t.hx = hx * inv_h;
    t.hy = hy * inv_h;
    t.hz = hz * inv_h;
    t.hdv = fmaxf(t.hx * p.vx + t.hy * p.vy + t.hz * p.vz, 0.0f);
    t.u = fminf(fmaxf(1.0f - t.hdv, 0.0f), 1.0f);
    const float u2 = t.u * t.u;
    t.P = u2 * u2 * t.u;
    // distribution_ggx
    const float a = p.rough * p.rough;
    t.a2 = a * a;
    t.ndh = fmaxf(p.nx * t.hx + p.ny * t.hy + p.nz * t.hz, 0.0f);
    t.den = t.ndh * t.ndh * (t.a2 - 1.0f) + 1.0f;
    t.Q = kPi * t.den * t.den + 0.0001f;
    t.D = t.a2 * hl_f_rcp(t.Q);
    // geometry_smith
    t.el = t.ndl * (1.0f - p.kk) + p.kk + 0.0001f;
    t.gl = t.ndl * hl_f_rcp(t.el);
    t.G = p.gv * t.gl;
    t.den2 = 4.0f * p.ndv * t.ndl + 0.0001f;
    const float spec = t.D * t.G * hl_f_rcp(t.den2);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        t.F[c] = p.f0[c] + (1.0f - p.f0[c]) * t.P;
        t.B[c] = (1.0f - t.F[c]) * p.one_m * p.alb[c] * kInvPi + spec * t.F[c];
    }
    return true;
}

// ------------------------------------------------------------------ kernel
template <int FA, int FN, int GRID, int MODE>
__global__ __launch_bounds__(256) void k_shade(ShadeArgs a) {
    __shared__ uint32_t s_list[4][kLdsCap];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t b = hl_xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t bx = b % a.block_tiles_x, by = b / a.block_tiles_x;
    const uint32_t x = bx * 32u + wave * 8u + (lane & 7u);
    const uint32_t y = by * 8u + (lane >> 3);
    const bool inside = x < a.width && y < a.height;
    const float depth = inside ? hl_load_depth(a.depth, x, y) : 1.0f;
    const bool valid = inside && depth < 1.0f;
    constexpr bool kWrites = MODE == kModeForward || MODE == kModeFusedL2;
    float4* out_px = (kWrites && inside) ? (float4*)(reinterpret_cast<uint8_t*>(a.out) + (size_t)y * a.out_pitch) + x : nullptr;
    if (kWrites && inside && !valid && (a.flags & HL_SHADE_CLEAR_SKY))
        hl_store_nt(out_px, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    if (hl_ballot(valid) == 0) return;

    // ---- per-pixel set-up (deferred_lighting.wgsl:999-1086)
    Pixel p;
    float4 A = make_float4(0, 0, 0, 0), Nr = A, O = A, E = A;
    if (valid) {
        A = hl_load_rgba<FA>(a.albedo, x, y);
        Nr = hl_load_rgba<FN>(a.normal, x, y);
        O = hl_load_rgba<FA>(a.orm, x, y);
        E = hl_load_rgba<FN>(a.emissive, x, y);
    } else {
        Nr = make_float4(0, 1, 0, 0);
    }
    {
        // N, V and NdV are per-pixel and exact: NdV cancels at grazing view
        // (dot of two unit vectors ~1e-6), where a last-bit difference in the
        // normalization becomes a percent-level difference in G1(NdV).
        const float3 n = hl_x_normalize(Nr.x, Nr.y, Nr.z);
        p.nx = n.x;
        p.ny = n.y;
        p.nz = n.z;
    }
    p.alb[0] = A.x;
    p.alb[1] = A.y;
    p.alb[2] = A.z;
    p.ao = O.x;
    p.rough = O.y;
    p.metal = O.z;
    p.one_m = 1.0f - p.metal;
    const GpuCameraUniforms* cam = a.cam;
    {
#pragma clang fp contract(off)
        const float u = ((float)x + 0.5f) / (float)a.width;
        const float v = ((float)y + 0.5f) / (float)a.height;
        const float nx = u * 2.0f - 1.0f, ny = 1.0f - v * 2.0f;
        const float* m = cam->inv_view_proj;
        const float hw = hl_x_mat_row(m, 3, nx, ny, depth, 1.0f);
        p.wx = hl_x_mat_row(m, 0, nx, ny, depth, 1.0f) / hw;
        p.wy = hl_x_mat_row(m, 1, nx, ny, depth, 1.0f) / hw;
        p.wz = hl_x_mat_row(m, 2, nx, ny, depth, 1.0f) / hw;
    }
    if (MODE == kModeBackward || MODE == kModeFusedL2) {
        // the differentiated forward resolves F0 from albedo and metallic
#pragma unroll
        for (int c = 0; c < 3; ++c)
            p.f0[c] = fminf(fmaxf(0.04f * (1.0f - p.metal) + p.alb[c] * p.metal, 0.0f), 0.999f);
    } else {
        p.f0[0] = fminf(fmaxf(Nr.w, 0.0f), 0.999f);
        p.f0[1] = fminf(fmaxf(O.w, 0.0f), 0.999f);
        p.f0[2] = fminf(fmaxf(E.w, 0.0f), 0.999f);
    }
    {
#pragma clang fp contract(off)
        const float3 v = hl_x_normalize(cam->position_near[0] - p.wx,
                                        cam->position_near[1] - p.wy,
                                        cam->position_near[2] - p.wz);
        p.vx = v.x;
        p.vy = v.y;
        p.vz = v.z;
        p.ndv = fmaxf(hl_x_dot3(p.nx, p.ny, p.nz, p.vx, p.vy, p.vz), 0.0f);
    }
    {
        const float r1 = p.rough + 1.0f;
        p.kk = r1 * r1 * 0.125f;
        p.gv = p.ndv * hl_f_rcp(p.ndv * (1.0f - p.kk) + p.kk + 0.0001f);
    }

    // ---- debug views (deferred_lighting.wgsl:1032-1038)
    if (MODE == kModeForward && a.debug_mode != 0u) {
        if (valid) {
            float4 o;
            if (a.debug_mode == 5u) {
                o = make_float4(p.nx * 0.5f + 0.5f, p.ny * 0.5f + 0.5f, p.nz * 0.5f + 0.5f, 1.0f);
            } else {
                o = make_float4(A.x, A.y, A.z, A.w);
            }
            hl_store_nt(out_px, o);
        }
        return;
    }

    // ---- cell of this pixel
    uint32_t cell = 0;
    if (GRID == HL_GRID_TILES) {
        cell = (y >> 4) * a.tile_stride_x + (x >> 4);
    } else {
        const uint32_t cx = x / a.cell_w, cy = y / a.cell_h;
        uint32_t cz = 0;
        if (valid && a.grid_z > 1u) {
            // guess from the view distance, then settle on the exact thresholds
            const float p22 = cam->proj[10], p32 = cam->proj[14];
            const float dist = p32 * hl_f_rcp(depth + p22);
            float g = __builtin_amdgcn_logf(fmaxf(dist, 1e-30f)) * a.log_scale + a.log_bias;
            int k = (int)floorf(g);
            k = k < 0 ? 0 : (k > (int)a.grid_z - 1 ? (int)a.grid_z - 1 : k);
            while (k > 0 && a.depth_split[k] > depth) --k;
            while (k < (int)a.grid_z - 1 && a.depth_split[k + 1] <= depth) ++k;
            cz = (uint32_t)k;
        }
        cell = cx + a.grid_x * (cy + a.grid_y * cz);
    }

    float lo[3] = {0.0f, 0.0f, 0.0f};
    uint32_t pairs = 0;

    uint64_t todo = hl_ballot(valid);
    while (todo) {
        const uint32_t leader = (uint32_t)__builtin_ctzll(todo);
        const uint32_t ucell = hl_readlane(cell, leader);
        const bool mine = valid && cell == ucell;
        todo &= ~hl_ballot(mine);
        // AABB of this cell's pixels in the wave
        const float inf = __builtin_inff();
        const float bx0 = hl_wave_min(mine ? p.wx : inf), bx1 = hl_wave_max(mine ? p.wx : -inf);
        const float by0 = hl_wave_min(mine ? p.wy : inf), by1 = hl_wave_max(mine ? p.wy : -inf);
        const float bz0 = hl_wave_min(mine ? p.wz : inf), bz1 = hl_wave_max(mine ? p.wz : -inf);
        // list of the cell
        uint32_t cnt, base;
        if (GRID == HL_GRID_TILES) {
            cnt = a.counts[ucell];
            base = ucell * HL_MAX_LIGHTS_PER_TILE;
        } else {
            cnt = a.counts[ucell];
            base = a.offsets[ucell];
        }
        cnt = hl_uniform(cnt);
        base = hl_uniform(base);
        if (base >= a.lists_capacity) cnt = 0;
        else if (cnt > a.lists_capacity - base) cnt = a.lists_capacity - base;
        const uint32_t* list = a.lists + base;

        uint32_t nl = 0;
        for (uint32_t j0 = 0; j0 < cnt; j0 += 64u) {
            const uint32_t j = j0 + lane;
            bool pass = false;
            uint32_t idx = 0;
            if (j < cnt) {
                idx = list[j];
                if (idx < a.num_lights) {
                    const float4 q = a.lights[idx].p0;
                    float dx = fmaxf(fmaxf(bx0 - q.x, q.x - bx1), 0.0f);
                    float dy = fmaxf(fmaxf(by0 - q.y, q.y - by1), 0.0f);
                    float dz = fmaxf(fmaxf(bz0 - q.z, q.z - bz1), 0.0f);
                    const float d2 = dx * dx + dy * dy + dz * dz;
                    pass = d2 <= q.w * q.w * 1.0001f + 1e-12f;
                }
            }
            const uint64_t m = hl_ballot(pass);
            if (pass) s_list[wave][nl + hl_prefix(m)] = idx;
            nl += (uint32_t)__popcll(m);
            if (nl > (uint32_t)kLdsCap - 64u || j0 + 64u >= cnt) {
                hl_wave_sync();
                for (uint32_t i = 0; i < nl; ++i) {
                    const uint32_t li = hl_uniform(s_list[wave][i]);
                    const HotLight L = a.lights[li];
                    if (mine) {
                        Term t;
                        bool in_range;
                        const bool act = eval_term<false>(L, p, t, in_range);
                        if (MODE == kModeCount) pairs += in_range ? 1u : 0u;
                        if (act) {
#pragma unroll
                            for (int c = 0; c < 3; ++c) lo[c] += t.B[c] * t.rad[c] * t.ndl;
                        }
                    }
                }
                hl_wave_sync();
                nl = 0;
            }
        }
    }

    if (MODE == kModeCount) {
        uint32_t tot = (uint32_t)hl_wave_sum((float)pairs);
        // exact for < 2^24 per wave (at most 64 pixels * lights per cell)
        if (lane == 0 && tot) atomicAdd(a.pair_counter, (unsigned long long)tot);
        return;
    }
    if (MODE == kModeForward && valid) {
        // ambient + combine (deferred_lighting.wgsl:1196-1231)
        const float ht = p.ny * 0.5f + 0.5f;
        float o[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float sky = a.sky[c];
            const float hemi = (sky * 0.15f * (1.0f - ht) + sky * ht) * p.alb[c];
            o[c] = lo[c] + hemi * p.ao + (c == 0 ? E.x : (c == 1 ? E.y : E.z));
        }
        hl_store_nt(out_px, make_float4(o[0], o[1], o[2], A.w));
    }
}

template <int FA, int FN, int MODE>
static hipError_t launch_grid(hipStream_t s, const ShadeArgs& a) {
    dim3 grid(a.num_blocks), block(256);
    if (a.grid_mode == HL_GRID_TILES)
        k_shade<FA, FN, HL_GRID_TILES, MODE><<<grid, block, 0, s>>>(a);
    else
        k_shade<FA, FN, HL_GRID_CLUSTERS, MODE><<<grid, block, 0, s>>>(a);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_fmt(hipStream_t s, int set, const ShadeArgs& a) {
    switch (set) {
        case kSetRef: return launch_grid<HL_FMT_RGBA8_UNORM, HL_FMT_RGBA16_FLOAT, MODE>(s, a);
        case kSet16: return launch_grid<HL_FMT_RGBA16_FLOAT, HL_FMT_RGBA16_FLOAT, MODE>(s, a);
        default: return launch_grid<HL_FMT_RGBA32_FLOAT, HL_FMT_RGBA32_FLOAT, MODE>(s, a);
    }
}

hipError_t launch_shade(hipStream_t s, int gbuf_set, int mode, const ShadeArgs& a) {
    if (a.num_blocks == 0) return hipSuccess;
    switch (mode) {
        case kModeForward: return launch_fmt<kModeForward>(s, gbuf_set, a);
        case kModeCount: return launch_fmt<kModeCount>(s, gbuf_set, a);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace hl
