// hl_reflect.hip -- the deferred pass's second draw, fs_reflection
// (crates/passes/3d/helio-pass-deferred-light/shaders/deferred_lighting.wgsl:
// 1266-1341): the indirect specular composite of the reflection captures'
// cube array (with the skylight in cube 0), screen-space and planar
// reflections and the radiance-cascade rough-specular fallback, added to the
// lit frame.
//
// One thread per pixel, a 16 x 16-pixel block of four 16 x 4 wavefronts.  A
// wave first finds, one capture per lane, the captures whose influence volume
// can reach the bounding box of its pixels' world positions (a conservative
// test), then walks only those, in order, with the records read by scalar
// loads; a lane leaves the walk once its coverage saturates.  The cube
// sampler is evaluated in FP32 exactly as the restatement defines it (face of
// the major axis, bilinear weights on uv * size - 0.5 clamped to the face's
// edge, linear mips), so that, with the IEEE division and square root, the
// composite follows the restatement operation for operation.
#include "hl_internal.hpp"
#include "hl_rc.hpp"

using namespace hl;

namespace {

constexpr uint32_t kMaxCaptures = 64;  // deferred_lighting.wgsl:656
constexpr float kEnvMaxLod = 8.0f;     // ENV_MAX_LOD, pbr_eval.wgsl:85

struct ReflectArgs {
    Tex normal, orm, emissive, depth;
    uint32_t width, height;
    const GpuCameraUniforms* cam;
    uint32_t enable_refl, enable_env, ncap, has_rc, debug;
    const GpuReflectionCapture* caps;
    const uint8_t* env;  // null: the black fallback cube array
    uint32_t env_fmt, env_face, env_mips, env_cubes;
    uint32_t cube_bytes;  // below 4 GiB
    Tex ssr, planar;  // data null: none
    uint32_t ssr_fmt, planar_fmt;
    Tex rc;  // data null: black
    uint32_t rc_fmt;
    float rc_min[3], rc_size[3];
    float* out;
    uint32_t out_pitch, flags;
};

__device__ __forceinline__ uint32_t texel_size(uint32_t fmt) {
    return fmt == HL_FMT_RGBA32_FLOAT ? 16u : fmt == HL_FMT_RGBA16_FLOAT ? 8u : 4u;
}

// RGBA texel at p in format fmt (uniform across the wave)
__device__ __forceinline__ float4 load_texel(const uint8_t* p, uint32_t fmt) {
    if (fmt == HL_FMT_RGBA32_FLOAT) return hl_ld4(reinterpret_cast<const float4*>(p));
    if (fmt == HL_FMT_RGBA16_FLOAT) return hl_decode<HL_FMT_RGBA16_FLOAT>(*hl_g(reinterpret_cast<const hl_u32x2*>(p)));
    return hl_decode<HL_FMT_RGBA8_UNORM>(*hl_g(reinterpret_cast<const uint32_t*>(p)));
}

// Where one sample's two mips sit within a cube, computed once per pixel
// (every cube of the array has the same layout; a cube is below 4 GiB).
struct MipPair {
    uint32_t off0, off1;
    uint32_t res0, res1;
    float t;
};

__device__ __forceinline__ MipPair mip_pair(const ReflectArgs& a, float lod) {
#pragma clang fp contract(off)
    MipPair mp;
    const float l = fminf(fmaxf(lod, 0.0f), (float)(a.env_mips - 1u));
    const float lf = floorf(l);
    const uint32_t m0 = (uint32_t)lf;
    mp.t = l - lf;
    const uint32_t tb = texel_size(a.env_fmt);
    uint32_t off = 0, r = a.env_face;
    for (uint32_t m = 0; m < m0; ++m) {
        off += 6u * r * r * tb;
        r = r > 1u ? r >> 1 : 1u;
    }
    mp.off0 = off;
    mp.res0 = r;
    if (m0 + 1u < a.env_mips) {
        mp.off1 = off + 6u * r * r * tb;
        mp.res1 = r > 1u ? r >> 1 : 1u;
    } else {
        mp.off1 = off;
        mp.res1 = r;
    }
    return mp;
}

__device__ __forceinline__ float3 cube_bilinear(const ReflectArgs& a, const uint8_t* mip, uint32_t res, uint32_t face,
                                                float u, float v) {
#pragma clang fp contract(off)
    const uint32_t tb = texel_size(a.env_fmt);
    const float fs = (float)res;
    const float x = u * fs - 0.5f, y = v * fs - 0.5f;
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const float hi = fs - 1.0f;
    const uint32_t ix0 = (uint32_t)fminf(fmaxf(x0, 0.0f), hi), ix1 = (uint32_t)fminf(fmaxf(x0 + 1.0f, 0.0f), hi);
    const uint32_t iy0 = (uint32_t)fminf(fmaxf(y0, 0.0f), hi), iy1 = (uint32_t)fminf(fmaxf(y0 + 1.0f, 0.0f), hi);
    const uint8_t* fb = mip + face * res * res * tb;
    const float4 t00 = load_texel(fb + (iy0 * res + ix0) * tb, a.env_fmt);
    const float4 t10 = load_texel(fb + (iy0 * res + ix1) * tb, a.env_fmt);
    const float4 t01 = load_texel(fb + (iy1 * res + ix0) * tb, a.env_fmt);
    const float4 t11 = load_texel(fb + (iy1 * res + ix1) * tb, a.env_fmt);
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    const float r = (t00.x * gx + t10.x * fx) * gy + (t01.x * gx + t11.x * fx) * fy;
    const float g = (t00.y * gx + t10.y * fx) * gy + (t01.y * gx + t11.y * fx) * fy;
    const float b = (t00.z * gx + t10.z * fx) * gy + (t01.z * gx + t11.z * fx) * fy;
    return make_float3(r, g, b);
}

// textureSampleLevel(env_cube, env_sampler, d, cube, lod).rgb (a.env non-null)
__device__ __forceinline__ float3 cube_sample(const ReflectArgs& a, const MipPair& mp, uint32_t cube, float dx,
                                              float dy, float dz) {
#pragma clang fp contract(off)
    cube = min(cube, a.env_cubes - 1u);
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    uint32_t face;
    float ma, sc, tc;
    if (ax >= ay && ax >= az) {
        face = dx >= 0.0f ? 0u : 1u;
        ma = ax;
        sc = dx >= 0.0f ? -dz : dz;
        tc = -dy;
    } else if (ay >= az) {
        face = dy >= 0.0f ? 2u : 3u;
        ma = ay;
        sc = dx;
        tc = dy >= 0.0f ? dz : -dz;
    } else {
        face = dz >= 0.0f ? 4u : 5u;
        ma = az;
        sc = dz >= 0.0f ? dx : -dx;
        tc = -dy;
    }
    const float u = (sc / ma + 1.0f) * 0.5f, v = (tc / ma + 1.0f) * 0.5f;
    const uint8_t* cb = a.env + (size_t)a.cube_bytes * cube;
    const float3 c0 = cube_bilinear(a, cb + mp.off0, mp.res0, face, u, v);
    const float3 c1 = cube_bilinear(a, cb + mp.off1, mp.res1, face, u, v);
    const float t = mp.t, s = 1.0f - t;
    return make_float3(c0.x * s + c1.x * t, c0.y * s + c1.y * t, c0.z * s + c1.z * t);
}

// capture_weight (:619-639)
__device__ __forceinline__ float capture_weight(const GpuReflectionCapture& cap, float px, float py, float pz) {
#pragma clang fp contract(off)
    if (cap.shape == HL_CAPTURE_SHAPE_BOX) {
        const float* m = cap.world_to_local;
        const float dx = cap.extents_transition[0] - fabsf(hl_x_mat_row(m, 0, px, py, pz, 1.0f));
        const float dy = cap.extents_transition[1] - fabsf(hl_x_mat_row(m, 1, px, py, pz, 1.0f));
        const float dz = cap.extents_transition[2] - fabsf(hl_x_mat_row(m, 2, px, py, pz, 1.0f));
        if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
        const float tr = fmaxf(cap.extents_transition[3], 1e-4f);
        return fminf(fmaxf(fminf(fminf(dx, dy), dz) / tr, 0.0f), 1.0f);
    }
    const float radius = cap.position_radius[3];
    const float vx = px - cap.position_radius[0], vy = py - cap.position_radius[1], vz = pz - cap.position_radius[2];
    const float dist = sqrtf(hl_x_dot3(vx, vy, vz, vx, vy, vz));
    if (dist > radius) return 0.0f;
    const float fade = fmaxf(radius * 0.1f, 1e-4f);
    return fminf(fmaxf((radius - dist) / fade, 0.0f), 1.0f);
}

// parallax_box (:599-615) or parallax_sphere (:577-591)
__device__ __forceinline__ float3 capture_dir(const GpuReflectionCapture& cap, float px, float py, float pz, float rx,
                                              float ry, float rz) {
#pragma clang fp contract(off)
    const float cx = cap.position_radius[0], cy = cap.position_radius[1], cz = cap.position_radius[2];
    float t;
    if (cap.shape == HL_CAPTURE_SHAPE_BOX) {
        const float* m = cap.world_to_local;
        float tv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float pl = hl_x_mat_row(m, k, px, py, pz, 1.0f);
            const float rl = hl_x_mat_row(m, k, rx, ry, rz, 0.0f);
            const float e = cap.extents_transition[k];
            const float inv = 1.0f / rl;
            tv[k] = fmaxf((-e - pl) * inv, (e - pl) * inv);
        }
        t = fminf(fminf(tv[0], tv[1]), tv[2]);
        if (t <= 0.0f || !(t < 3.402823e38f)) return make_float3(rx, ry, rz);
    } else {
        const float ox = px - cx, oy = py - cy, oz = pz - cz;
        const float b = hl_x_dot3(ox, oy, oz, rx, ry, rz);
        const float r = cap.position_radius[3];
        const float c = hl_x_dot3(ox, oy, oz, ox, oy, oz) - r * r;
        const float disc = b * b - c;
        if (disc < 0.0f) return make_float3(rx, ry, rz);
        t = -b + sqrtf(disc);
        if (t <= 0.0f) return make_float3(rx, ry, rz);
    }
    return make_float3((px + rx * t) - cx, (py + ry * t) - cy, (pz + rz * t) - cz);
}

// Whether capture cap can give any position of the box [lo, hi] a positive
// capture_weight, with a margin far above the rounding of both this test and
// the per-pixel weight (non-finite fields keep the capture).
__device__ __forceinline__ bool capture_may_touch(const GpuReflectionCapture& cap, const float lo[3],
                                                  const float hi[3]) {
    if (cap.shape == HL_CAPTURE_SHAPE_BOX) {
        // the box's positions map into the local box centre +- |A| half-extent
        const float* m = cap.world_to_local;
        float c[3], h[3], mc[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c[k] = (lo[k] + hi[k]) * 0.5f;
            h[k] = (hi[k] - lo[k]) * 0.5f;
            mc[k] = fabsf(c[k]);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float cl = hl_x_mat_row(m, k, c[0], c[1], c[2], 1.0f);
            const float a0 = fabsf(m[0 * 4 + k]), a1 = fabsf(m[1 * 4 + k]), a2 = fabsf(m[2 * 4 + k]);
            const float hl = a0 * h[0] + a1 * h[1] + a2 * h[2];
            const float mag = a0 * mc[0] + a1 * mc[1] + a2 * mc[2] + fabsf(m[3 * 4 + k]) + hl;
            const float e = cap.extents_transition[k];
            if (fabsf(cl) - hl > e + 1e-5f * (mag + fabsf(e)) + 1e-6f) return false;
        }
        return true;
    }
    float d2 = 0.0f, mag = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float c = cap.position_radius[k];
        const float d = fmaxf(fmaxf(lo[k] - c, c - hi[k]), 0.0f);
        d2 += d * d;
        mag = fmaxf(mag, fmaxf(fabsf(c), fmaxf(fabsf(lo[k]), fabsf(hi[k]))));
    }
    const float r = cap.position_radius[3];
    const float rr = r + 1e-5f * (mag + fabsf(r)) + 1e-6f;
    return !(d2 > rr * rr);
}

// Captures (bit i: capture i) that may weigh on a live pixel of the wave:
// every capture that gives a live pixel a positive weight is set (all of them
// when a live position is not finite).  Called by every lane of the wave.
__device__ __forceinline__ uint64_t wave_captures(const ReflectArgs& a, bool live, float px, float py, float pz) {
    const uint64_t all = a.ncap >= 64u ? ~0ull : ((1ull << a.ncap) - 1ull);
    const bool finite = __builtin_isfinite(px) && __builtin_isfinite(py) && __builtin_isfinite(pz);
    if (hl_ballot(live && !finite) != 0) return all;
    const float inf = __builtin_inff();
    float lo[3] = {live ? px : inf, live ? py : inf, live ? pz : inf};
    float hi[3] = {live ? px : -inf, live ? py : -inf, live ? pz : -inf};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = hl_uniform(hl_wave_min(lo[k]));
        hi[k] = hl_uniform(hl_wave_max(hi[k]));
    }
    const uint32_t j = hl_lane();
    bool hit = false;
    if (j < a.ncap) {
        const GpuReflectionCapture& cap = a.caps[j];
        hit = cap.cubemap_index >= 0 && capture_may_touch(cap, lo, hi);
    }
    return hl_ballot(hit);
}

// sample_reflection_environment (:646-691) over the captures of mask (a
// capture outside it weighs 0 on every pixel of the wave, which the shader's
// loop skips); the caller checks a.env and enable_reflections
__device__ __forceinline__ float3 reflection_env(const ReflectArgs& a, const MipPair& mp, uint64_t mask, float px,
                                                 float py, float pz, float rx, float ry, float rz) {
#pragma clang fp contract(off)
    float ar = 0.0f, ag = 0.0f, ab = 0.0f, aa = 0.0f;
    for (uint64_t bits = mask; bits; bits &= bits - 1ull) {
        const uint32_t i = (uint32_t)__builtin_ctzll(bits);
        if (aa >= 0.999f) break;
        const GpuReflectionCapture& cap = a.caps[i];
        const float w = capture_weight(cap, px, py, pz);
        if (w <= 0.0f) continue;
        const float3 d = capture_dir(cap, px, py, pz, rx, ry, rz);
        const float3 s = cube_sample(a, mp, (uint32_t)cap.cubemap_index, d.x, d.y, d.z);
        const float br = cap.brightness;
        const float contrib = w * (1.0f - aa);
        ar = ar + (s.x * br) * contrib;
        ag = ag + (s.y * br) * contrib;
        ab = ab + (s.z * br) * contrib;
        aa = aa + contrib;
    }
    if (aa < 0.999f) {
        const float3 sky = cube_sample(a, mp, 0u, rx, ry, rz);
        const float rest = 1.0f - aa;
        ar = ar + sky.x * rest;
        ag = ag + sky.y * rest;
        ab = ab + sky.z * rest;
    }
    return make_float3(ar, ag, ab);
}

__device__ __forceinline__ bool debug_adds_nothing(uint32_t m) {
    return m == 1u || m == 2u || m == 4u || m == 5u || m == 10u || m == 11u || m == 20u || m == 21u;
}

template <int FA, int FN>
__global__ __launch_bounds__(256) void k_reflection(ReflectArgs a) {
#pragma clang fp contract(off)
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t x = blockIdx.x * 16u + (lane & 15u);
    const uint32_t y = blockIdx.y * 16u + wave * 4u + (lane >> 4);
    // lanes stay active until the wave's capture set is known; sky pixels
    // (depth >= 1) are discarded, those outside the frame have nothing to do
    const bool inside = x < a.width && y < a.height;
    const float depth = inside ? hl_load_depth(a.depth, x, y) : 1.0f;
    const bool live = inside && !(depth >= 1.0f);
    if (hl_ballot(live) == 0) return;
    float4* o = reinterpret_cast<float4*>(reinterpret_cast<uint8_t*>(a.out) + (size_t)y * a.out_pitch) + x;
    float4 ssr = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (live && a.ssr.data)
        ssr = load_texel(a.ssr.data + (size_t)y * a.ssr.pitch + (size_t)x * texel_size(a.ssr_fmt), a.ssr_fmt);
    if (a.debug == 30u) {
        if (live) hl_store_nt(o, make_float4(ssr.w, ssr.w, ssr.w, 1.0f));
        return;
    }
    if (a.debug == 31u) {
        if (live) hl_store_nt(o, make_float4(ssr.x, ssr.y, ssr.z, 1.0f));
        return;
    }
    if (debug_adds_nothing(a.debug)) {
        if (live && (a.flags & HL_REFLECTION_STORE)) hl_store_nt(o, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        return;
    }
    typename TexRaw<FN>::type rN{}, rE{};
    typename TexRaw<FA>::type rO{};
    if (live) {
        rN = hl_fetch<FN>(a.normal, x, y);
        rO = hl_fetch<FA>(a.orm, x, y);
        rE = hl_fetch<FN>(a.emissive, x, y);
        if (a.planar.data)
            pl = load_texel(a.planar.data + (size_t)y * a.planar.pitch + (size_t)x * texel_size(a.planar_fmt),
                            a.planar_fmt);
    }
    const float4 Nr = hl_decode<FN>(rN), O = hl_decode<FA>(rO), E = hl_decode<FN>(rE);
    const float3 n = hl_x_normalize(Nr.x, Nr.y, Nr.z);
    const float rough = O.y;
    const float f0[3] = {fminf(fmaxf(Nr.w, 0.0f), 0.999f), fminf(fmaxf(O.w, 0.0f), 0.999f),
                         fminf(fmaxf(E.w, 0.0f), 0.999f)};
    const float ao = O.x * 1.0f;  // screen-space AO: its white fallback
    // world position as fs_reflection reconstructs it (:1298-1301)
    const float u = ((float)x + 0.5f) / (float)a.width;
    const float v = ((float)y + 0.5f) / (float)a.height;
    const float ndx = u * 2.0f - 1.0f, ndy = 1.0f - v * 2.0f;
    const float* m = a.cam->inv_view_proj;
    const float hw = hl_x_mat_row(m, 3, ndx, ndy, depth, 1.0f);
    const float px = hl_x_mat_row(m, 0, ndx, ndy, depth, 1.0f) / hw;
    const float py = hl_x_mat_row(m, 1, ndx, ndy, depth, 1.0f) / hw;
    const float pz = hl_x_mat_row(m, 2, ndx, ndy, depth, 1.0f) / hw;
    const float3 V = hl_x_normalize(a.cam->position_near[0] - px, a.cam->position_near[1] - py,
                                    a.cam->position_near[2] - pz);
    const float ndv = fmaxf(hl_x_dot3(n.x, n.y, n.z, V.x, V.y, V.z), 0.0f);
    // fresnel_schlick_roughness (pbr_eval.wgsl:63-66)
    const float uu = fminf(fmaxf(1.0f - ndv, 0.0f), 1.0f);
    const float u2 = uu * uu;
    const float p5 = u2 * u2 * uu;
    float fi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) fi[k] = f0[k] + (fmaxf(1.0f - rough, f0[k]) - f0[k]) * p5;
    // R = reflect(-V, N)
    const float ix = -V.x, iy = -V.y, iz = -V.z;
    const float dni = hl_x_dot3(n.x, n.y, n.z, ix, iy, iz);
    const float rx = ix - 2.0f * dni * n.x, ry = iy - 2.0f * dni * n.y, rz = iz - 2.0f * dni * n.z;
    const bool use_env = a.enable_env && a.env && a.enable_refl;
    const uint64_t capmask = use_env && a.ncap ? wave_captures(a, live, px, py, pz) : 0ull;
    if (!live) return;
    float spec[3] = {0.0f, 0.0f, 0.0f};
    if (a.enable_env) {
        // a black cube array samples 0 everywhere: the environment term is 0
        float3 env = make_float3(0.0f, 0.0f, 0.0f);
        if (use_env) env = reflection_env(a, mip_pair(a, rough * kEnvMaxLod), capmask, px, py, pz, rx, ry, rz);
        // env_brdf_approx (pbr_eval.wgsl:70-76)
        const float r0 = rough * -1.0f + 1.0f, r1 = rough * -0.0275f + 0.0425f;
        const float r2 = rough * -0.572f + 1.04f, r3 = rough * 0.022f + -0.04f;
        const float a004 = fminf(r0 * r0, exp2f(-9.28f * ndv)) * r0 + r1;
        const float bx = -1.04f * a004 + r2, by = 1.04f * a004 + r3;
        const float e3[3] = {env.x, env.y, env.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) spec[k] = e3[k] * (f0[k] * bx + by);
    }
    const float s4[3] = {ssr.x, ssr.y, ssr.z}, p4[3] = {pl.x, pl.y, pl.z};
    if (a.enable_refl && ssr.w > 0.0f) {
#pragma unroll
        for (int k = 0; k < 3; ++k) spec[k] = spec[k] * (1.0f - ssr.w) + (s4[k] * fi[k]) * ssr.w;
    }
    if (a.enable_refl && pl.w > 0.0f) {
#pragma unroll
        for (int k = 0; k < 3; ++k) spec[k] = spec[k] * (1.0f - pl.w) + (p4[k] * fi[k]) * pl.w;
    }
    if (a.has_rc && rough > 0.6f) {
        // sample_rc_specular (:929-983) along normalize(mix(R, N, roughness * 0.5))
        float3 rc = make_float3(0.0f, 0.0f, 0.0f);
        if (a.rc.data) {
            const float t = rough * 0.5f, s = 1.0f - t;
            const float3 L = hl_x_normalize(rx * s + n.x * t, ry * s + n.y * t, rz * s + n.z * t);
            rc = rc_irradiance(a, px, py, pz, L.x, L.y, L.z);
        }
        const float w = rc_smoothstep(0.6f, 0.9f, rough) * 0.4f;
        const float c3[3] = {rc.x, rc.y, rc.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) spec[k] = spec[k] * (1.0f - w) + c3[k] * w;
    }
    if (a.flags & HL_REFLECTION_STORE) {
        hl_store_nt(o, make_float4(spec[0] * ao, spec[1] * ao, spec[2] * ao, 0.0f));
    } else {
        const float4 c = hl_ld4(o);
        hl_store_nt(o, make_float4(c.x + spec[0] * ao, c.y + spec[1] * ao, c.z + spec[2] * ao, c.w));
    }
}

template <int FA, int FN>
hipError_t launch(hipStream_t s, const ReflectArgs& a) {
    const dim3 grid((a.width + 15u) / 16u, (a.height + 15u) / 16u);
    k_reflection<FA, FN><<<grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

uint32_t host_texel_size(uint32_t fmt) {
    return fmt == HL_FMT_RGBA32_FLOAT ? 16u : fmt == HL_FMT_RGBA16_FLOAT ? 8u : fmt == HL_FMT_RGBA8_UNORM ? 4u : 0u;
}

// an optional full-resolution RGBA16F/RGBA32F input
hl_status check_screen_tex(hl_context* c, const hl_texture& t, uint32_t width, const char* name) {
    if (!t.data_d) return HL_OK;
    if (t.format != HL_FMT_RGBA16_FLOAT && t.format != HL_FMT_RGBA32_FLOAT)
        return fail(c, HL_EINVAL, "%s must be RGBA16_FLOAT or RGBA32_FLOAT", name);
    const uint32_t tb = host_texel_size(t.format);
    if ((uint64_t)t.pitch_bytes < (uint64_t)width * tb || t.pitch_bytes % tb || ((uintptr_t)t.data_d) % tb)
        return fail(c, HL_EINVAL, "%s: pitch %u invalid for width %u or data not aligned", name, t.pitch_bytes,
                    width);
    return HL_OK;
}

}  // namespace

extern "C" {

HL_API hl_status hl_deferred_reflection(hl_context* c, const GpuCameraUniforms* camera_d, const DeferredGlobals* g,
                                        const hl_gbuffer* gb, const hl_reflection_inputs* in, float* out_d,
                                        uint32_t out_pitch_bytes, uint32_t flags) {
    if (!c) return HL_EINVAL;
    if (!camera_d || !g || !in || !out_d) return fail(c, HL_EINVAL, "camera, globals, inputs or output is null");
    if (flags & ~(uint32_t)HL_REFLECTION_STORE) return fail(c, HL_EINVAL, "unknown reflection flags 0x%x", flags);
    const uint32_t dm = g->debug_mode;
    if (!(dm == 0u || dm == 30u || dm == 31u || dm == 1u || dm == 2u || dm == 4u || dm == 5u || dm == 10u ||
          dm == 11u || dm == 20u || dm == 21u))
        return fail(c, HL_EUNSUPPORTED, "debug_mode %u is not supported by the reflection composite", dm);
    int set = 0;
    hl_status s = check_gbuffer(c, gb, &set);
    if (s != HL_OK) return s;
    if (out_pitch_bytes < 16u * gb->width || out_pitch_bytes % 16 != 0 || ((uintptr_t)out_d) % 16)
        return fail(c, HL_EINVAL, "output pitch %u invalid for width %u (RGBA32F, 16-B aligned)", out_pitch_bytes,
                    gb->width);
    ReflectArgs a{};
    a.normal = Tex{(const uint8_t*)gb->normal.data_d, gb->normal.pitch_bytes};
    a.orm = Tex{(const uint8_t*)gb->orm.data_d, gb->orm.pitch_bytes};
    a.emissive = Tex{(const uint8_t*)gb->emissive.data_d, gb->emissive.pitch_bytes};
    a.depth = Tex{(const uint8_t*)gb->depth.data_d, gb->depth.pitch_bytes};
    a.width = gb->width;
    a.height = gb->height;
    a.cam = camera_d;
    a.enable_refl = g->enable_reflections != 0u;
    a.enable_env = g->enable_env_reflections != 0u;
    a.debug = dm;
    a.has_rc = g->has_rc_gi > 0u;
    a.out = out_d;
    a.out_pitch = out_pitch_bytes;
    a.flags = flags;
    const hl_cube_array& e = in->env;
    if (e.data_d) {
        const uint32_t tb = host_texel_size(e.format);
        if (!tb) return fail(c, HL_EINVAL, "env cube array format %u unsupported", e.format);
        if (e.face_size == 0 || e.face_size > 16384)
            return fail(c, HL_EINVAL, "env face_size %u out of range (1..16384)", e.face_size);
        uint32_t max_mips = 1;
        while ((e.face_size >> max_mips) != 0u) ++max_mips;
        if (e.mip_levels == 0 || e.mip_levels > max_mips)
            return fail(c, HL_EINVAL, "env mip_levels %u out of range (1..%u)", e.mip_levels, max_mips);
        if (e.num_cubes == 0) return fail(c, HL_EINVAL, "env cube array has no cubes");
        if (((uintptr_t)e.data_d) % tb) return fail(c, HL_EINVAL, "env data not aligned to %u bytes", tb);
        size_t cb = 0;
        for (uint32_t m = 0; m < e.mip_levels; ++m) {
            const size_t r = (e.face_size >> m) ? (e.face_size >> m) : 1u;
            cb += 6u * r * r * tb;
        }
        if (cb > 0xFFFFFFFFull)
            return fail(c, HL_EUNSUPPORTED, "one cube of the env array takes %zu bytes (at most 4 GiB)", cb);
        a.env = (const uint8_t*)e.data_d;
        a.env_fmt = e.format;
        a.env_face = e.face_size;
        a.env_mips = e.mip_levels;
        a.env_cubes = e.num_cubes;
        a.cube_bytes = (uint32_t)cb;
        a.ncap = g->reflection_capture_count < kMaxCaptures ? g->reflection_capture_count : kMaxCaptures;
        if (a.enable_env && a.enable_refl && a.ncap && !in->captures_d)
            return fail(c, HL_EINVAL, "%u reflection captures but captures_d is null", a.ncap);
        if (((uintptr_t)in->captures_d) % 16) return fail(c, HL_EINVAL, "captures_d not 16-byte aligned");
        a.caps = in->captures_d;
    }
    if ((s = check_screen_tex(c, in->ssr, gb->width, "ssr")) != HL_OK) return s;
    if ((s = check_screen_tex(c, in->planar, gb->width, "planar")) != HL_OK) return s;
    a.ssr = Tex{(const uint8_t*)in->ssr.data_d, in->ssr.pitch_bytes};
    a.ssr_fmt = in->ssr.format;
    a.planar = Tex{(const uint8_t*)in->planar.data_d, in->planar.pitch_bytes};
    a.planar_fmt = in->planar.format;
    const hl_texture& rc = in->rc_cascade0;
    if (rc.data_d && a.has_rc) {
        if (rc.format != HL_FMT_RGBA16_FLOAT && rc.format != HL_FMT_RGBA32_FLOAT)
            return fail(c, HL_EINVAL, "rc_cascade0 must be RGBA16_FLOAT or RGBA32_FLOAT");
        const uint32_t tb = host_texel_size(rc.format);
        if (rc.pitch_bytes < 64u * tb || rc.pitch_bytes % 16u || ((uintptr_t)rc.data_d) % 16u)
            return fail(c, HL_EINVAL, "rc_cascade0: pitch %u invalid for 64 texels or not 16-byte aligned",
                        rc.pitch_bytes);
        a.rc = Tex{(const uint8_t*)rc.data_d, rc.pitch_bytes};
        a.rc_fmt = rc.format;
        for (int k = 0; k < 3; ++k) {
            a.rc_min[k] = g->rc_world_min[k];
            a.rc_size[k] = g->rc_world_max[k] - g->rc_world_min[k];
        }
    }
    HL_HIP(c, hipSetDevice(c->device));
    hipError_t err;
    if (set == kSetRef) err = launch<HL_FMT_RGBA8_UNORM, HL_FMT_RGBA16_FLOAT>(c->stream, a);
    else if (set == kSet16) err = launch<HL_FMT_RGBA16_FLOAT, HL_FMT_RGBA16_FLOAT>(c->stream, a);
    else err = launch<HL_FMT_RGBA32_FLOAT, HL_FMT_RGBA32_FLOAT>(c->stream, a);
    if (err != hipSuccess) return hip_fail(c, err, "reflection launch");
    return HL_OK;
}

}  // extern "C"
