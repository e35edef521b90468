// hl_device.hpp -- device-side building blocks shared by the gfx950 kernels.
//
// Two kinds of arithmetic live here:
//  * "exact" functions (hl_x_*) compile with FP contraction off and use the
//    correctly rounded IEEE division and square root (the HIP default), one
//    binary32 operation per source operation, in the order light_cull.wgsl and
//    deferred_lighting.wgsl are restated in.  Everything that decides list
//    membership, the pixel-to-cell mapping, the reconstructed world position and
//    the all(radiance < 0.002) cut-off goes through them, so those decisions
//    are identical to an IEEE FP32 evaluation of the shaders.
//  * "fast" functions (hl_f_*) use contraction and the hardware rcp/rsq/sqrt
//    for the BRDF arithmetic, whose result only needs 1e-4 relative accuracy.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/helio_hip/helio_hip.h"

#define HL_WAVE 64

namespace hl {

// ------------------------------------------------------------------ records

// Per-light record the shading kernels read (80 B).  Built from GpuLight by
// the prep kernel each call:
//   p0 = (position.xyz, range)          range = +inf for directional lights
//   p1 = (color.rgb * intensity, cos inner)
//   p2 = (direction.xyz, cos outer)     cos outer = -2 for non-spot lights
//   p3 = (1 / range^2, intensity, cone margin, shadow word as bits)
//   p4 = (lo, hi, ci, 0) with ci = max(p1.xyz): the band of the first
//        per-pixel decision on the squared distance d2 (d2 < lo: in, d2 > hi:
//        the term is zero, otherwise decided exactly) -- the range test alone
//        for the forward-lit shader, the range test and the 0.002 cut of the
//        deferred shader together (make_hot_light)
// The cone margin, 1e-5 * |direction|_1, bounds (tenfold) the difference
// between a fast cos(angle) and the exact one of deferred_lighting.wgsl:739.
struct alignas(16) HotLight {
    float4 p0, p1, p2, p3, p4;
};

// Per-light record the culling kernels read (64 B):
//   vpos = (view-space position, range)  (range = +inf: directional, always listed)
//   cone = (view-space direction (not normalized), cos outer)  (cos outer = -2: no cone)
//   axis = (unit view-space axis, cos of the outer half-angle clamped to 1),
//   aux  = (sin of that angle, 0, 0, 0): the light-only terms of the cone
//          test, computed once by the prep kernel; a spot light whose cone
//          cannot reject anything (zero axis or half-angle >= 90 degrees)
//          has cone.w = -2 like the other lights
struct alignas(16) CullLight {
    float4 vpos, cone, axis, aux;
};

constexpr float kNoCone = -2.0f;

// Shadow word of a shading record (p3.w as bits): the light's first shadow
// layer (28 bits; a larger shadow_index saturates, past any atlas) and its
// kind in the top bits (0 directional, 1 point, 2 any other type: the spot
// path of shadow_factor); all ones when the light casts no shadow.  Only
// lights below MAX_SHADOW_LIGHTS = 42 (deferred_lighting.wgsl:20) carry one.
constexpr uint32_t kMaxShadowLights = 42u;
constexpr uint32_t kNoShadowWord = 0xFFFFFFFFu;
__host__ __device__ inline uint32_t shadow_word(uint32_t index, uint32_t type, uint32_t shadow_index) {
    if (index >= kMaxShadowLights || shadow_index == HL_NO_SHADOW) return kNoShadowWord;
    const uint32_t kind = type == HL_LIGHT_DIRECTIONAL ? 0u : (type == HL_LIGHT_POINT ? 1u : 2u);
    return (shadow_index < 0x0FFFFFFFu ? shadow_index : 0x0FFFFFFFu) | kind << 28;
}

// The shadow inputs of one shading launch (hl_shadow_maps): the two atlases,
// the light matrices and the ShadowConfig fields shadow_factor reads.
struct ShadowView {
    const float* atlas;
    const float* static_atlas;
    const float* matrices;  // 16 per layer
    const float* vogel;     // Vogel disk tables (k_vogel_table)
    uint32_t layers, size;
    uint32_t pcf_count, enable_pcss, blocker_samples, filter_samples;
    float filter_radius[4], light_size[4];
    float csm[4];
};

// Band (lo, hi) of the first per-pixel decision on d2 = |light - pixel|^2
// for a light of the given range (squared, range2) and largest radiance
// channel ci.  The range test alone: r^2 (1 -+ 1e-5), beyond the rounding of
// either side.  With the cut (deferred shading), d2 must also pass
// all(radiance < 0.002) of pbr_direct_light (deferred_lighting.wgsl:781) for
// an unshadowed point light, whose largest channel ci (1 - (d2/r^2)^2) /
// (d2 + 1e-4) decreases in d2 and crosses 0.002 at d2* = x r^2, x the
// positive root of ci x^2 + 0.002 r^2 x - (ci - 2e-7) = 0 (computed in FP64
// in the stable form).  The band around d2* is +-1e-4 relative: there the
// function moves by at least ~1e-4 relative, far beyond the rounding of its
// FP32 evaluation (a few ulp, and the 1 - nd^4 cancellation near the range
// is steeper still); below d2* = 1e-2 the band widens to [0, 2 d2* + 1e-3].
// The smaller of the two lower and of the two upper bounds is kept; lo is
// rounded down and hi up to FP32.
__device__ inline float2 decision_band(float range, float ci, bool cut) {
    const double R = (double)range * (double)range;
    double lo = R * (1.0 - 1e-5), hi = R * (1.0 + 1e-5);
    if (cut) {
        double tlo = 0.0, thi = 1e-3;  // ci <= 2e-7: at most 0.002 at d2 = 0
        const double c = (double)ci;
        if (c > 2e-7) {
            const double b = 0.002 * R, cc = c - 2e-7;
            const double d2c = R * (2.0 * cc / (b + sqrt(b * b + 4.0 * c * cc)));
            if (d2c < 1e-2) {
                tlo = 0.0;
                thi = 2.0 * d2c + 1e-3;
            } else {
                tlo = d2c * (1.0 - 1e-4);
                thi = d2c * (1.0 + 1e-4);
            }
        }
        lo = fmin(lo, tlo);
        hi = fmin(hi, thi);
    }
    return make_float2(__double2float_rd(lo), __double2float_ru(hi));
}

// The shading record of one light from its fields: position (px, py, pz),
// direction (dx, dy, dz) as stored (a spot light's axis, the direction a
// directional light shines along), colour and intensity ci, range, the cosines
// of the outer and inner cone angles, and the light type; cut: the record's
// band includes the 0.002 cut (deferred shading) or not (forward-lit).
__device__ __forceinline__ HotLight make_hot_light(float px, float py, float pz, float dx, float dy, float dz,
                                                   float4 ci, float range_field, float outer, float inner,
                                                   uint32_t type, uint32_t shadow, bool cut) {
    const float range = type == HL_LIGHT_DIRECTIONAL ? __builtin_inff() : range_field;
    HotLight h;
    h.p0 = make_float4(px, py, pz, range);
    {
#pragma clang fp contract(off)
        h.p1 = make_float4(ci.x * ci.w, ci.y * ci.w, ci.z * ci.w, inner);
    }
    h.p2 = make_float4(dx, dy, dz, type == HL_LIGHT_SPOT ? outer : kNoCone);
    const float margin = 1e-5f * (fabsf(dx) + fabsf(dy) + fabsf(dz));
    h.p3 = make_float4(1.0f / (range * range), ci.w, margin, __uint_as_float(shadow));
    const float cmax = fmaxf(fmaxf(h.p1.x, h.p1.y), h.p1.z);
    const float2 band = decision_band(range, cmax, cut);
    h.p4 = make_float4(band.x, band.y, cmax, 0.0f);
    return h;
}

// Texel formats the shading kernels are instantiated for.
enum GbufSet : int { kSetRef = 0, kSet16 = 1, kSet32 = 2 };

struct Tex {
    const uint8_t* data;
    uint32_t pitch;
};

// Everything a shading launch needs, passed by value as a kernel argument.
struct ShadeArgs {
    // camera (copied from device memory by the prep kernel into this form)
    const GpuCameraUniforms* cam;
    uint32_t width, height;
    float rcp_w, rcp_h;       // correctly rounded 1 / width, 1 / height
    Tex albedo, normal, orm, emissive, depth;
    const HotLight* lights;
    uint32_t num_lights;
    // grid
    uint32_t grid_mode;     // HL_GRID_TILES / HL_GRID_CLUSTERS
    uint32_t grid_x, grid_y, grid_z, cell_w, cell_h;
    const uint32_t* lists;
    const uint32_t* counts;
    const uint32_t* offsets;
    uint32_t lists_capacity;  // entries readable in lists
    uint32_t tile_stride_x;   // tiles: DeferredGlobals.num_tiles_x
    float log_scale, log_bias;// slice guess: floor(log2(z)*scale + bias), fixed up by depth_split
    float depth_split[HL_MAX_SLICES + 1];  // hl_cluster_params.depth_split
    // globals
    float sky[3];
    uint32_t debug_mode;
    uint32_t flags;
    // outputs
    float* out;
    uint32_t out_pitch;
    unsigned long long* pair_counter;
    // adjoint
    const float* dout;       // dL/dout RGBA32F (or target for the fused L2 variant)
    uint32_t dout_pitch;
    float* d_pixel;          // 5 planes
    double* d_light;         // kGradSlots x 4 per light: dpos.xyz, dI (FP64 accumulation)
    double* loss;            // fused L2 loss accumulator
    uint32_t block_tiles_x;  // block tiles per row
    uint32_t num_blocks;     // block tiles of the view
    uint32_t tw_log2;        // wavefront tile: (1 << tw_log2) x (64 >> tw_log2) pixels
    uint32_t out_f16;        // the output is RGBA16F (8 B per pixel, IEEE round to nearest even)
    ShadowView shadow;       // kModeExtended only (layers 0: no shadows)
    // kModeExtended: radiance-cascade probe atlas (rc_fmt 0: no RC GI) and
    // its world box (min, max - min)
    Tex rc;
    uint32_t rc_fmt;
    float rc_min[3], rc_size[3];
    // the probe atlas sampled per pixel ahead of the launch (k_rc_irradiance:
    // width * height RGBA32F, rgb = sample_rc_irradiance at the pixel)
    const float4* rc_irr;
};

// One view of a batched persistent launch (device memory).
struct BatchShadeView {
    ShadeArgs sa;
    uint32_t mode;  // kModeForward or kModeFusedL2
    uint32_t _pad[3];
};

// ------------------------------------------------------------- exact math
__device__ __forceinline__ float hl_x_dot3(float ax, float ay, float az, float bx, float by,
                                           float bz) {
#pragma clang fp contract(off)
    float s = ax * bx;
    s = s + ay * by;
    s = s + az * bz;
    return s;
}

// M (column-major) * (x, y, z, w), one component, columns accumulated in order.
__device__ __forceinline__ float hl_x_mat_row(const float* m, int row, float x, float y, float z,
                                              float w) {
#pragma clang fp contract(off)
    float s = m[0 * 4 + row] * x;
    s = s + m[1 * 4 + row] * y;
    s = s + m[2 * 4 + row] * z;
    s = s + m[3 * 4 + row] * w;
    return s;
}

__device__ __forceinline__ float3 hl_x_normalize(float x, float y, float z) {
#pragma clang fp contract(off)
    float l = sqrtf(hl_x_dot3(x, y, z, x, y, z));
    return make_float3(x / l, y / l, z / l);
}

// light_cull.wgsl:85-100 make_plane_from_ndc_edge (d = 0 dropped; the caller adds 0.0f).
__device__ __forceinline__ float3 hl_x_plane(float p00, float p11, float p0x, float p0y,
                                             float p1x, float p1y) {
#pragma clang fp contract(off)
    float3 r0 = hl_x_normalize(p0x / p00, p0y / p11, -1.0f);
    float3 r1 = hl_x_normalize(p1x / p00, p1y / p11, -1.0f);
    float cx = r0.y * r1.z - r0.z * r1.y;
    float cy = r0.z * r1.x - r0.x * r1.z;
    float cz = r0.x * r1.y - r0.y * r1.x;
    return hl_x_normalize(cx, cy, cz);
}

// light_cull.wgsl:105-117 sphere_inside_tile_frustum for one plane.
__device__ __forceinline__ bool hl_x_plane_pass(float3 n, float4 c, float r) {
#pragma clang fp contract(off)
    float d = hl_x_dot3(n.x, n.y, n.z, c.x, c.y, c.z) + 0.0f;
    return !(d < -r);
}

__device__ __forceinline__ bool hl_x_sphere_aabb(float4 c, float r, float3 lo, float3 hi) {
#pragma clang fp contract(off)
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    if (c.x < lo.x) dx = lo.x - c.x; else if (c.x > hi.x) dx = c.x - hi.x;
    if (c.y < lo.y) dy = lo.y - c.y; else if (c.y > hi.y) dy = c.y - hi.y;
    if (c.z < lo.z) dz = lo.z - c.z; else if (c.z > hi.z) dz = c.z - hi.z;
    float d2 = 0.0f;
    d2 = d2 + dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    return d2 <= r * r;
}

// Cone against the AABB's bounding sphere (see the cluster membership rule),
// in two parts.  The light part (axis length, unit axis, cos and sin of the
// half-angle) returns false when the cone cannot reject anything.
__device__ __forceinline__ bool hl_x_cone_light(float4 cone, float4* axis, float* sin_t) {
#pragma clang fp contract(off)
    float alen = sqrtf(hl_x_dot3(cone.x, cone.y, cone.z, cone.x, cone.y, cone.z));
    if (!(alen > 0.0f)) return false;
    float cos_t = cone.w / alen;
    if (!(cos_t > 0.0f)) return false;
    if (cos_t > 1.0f) cos_t = 1.0f;
    *sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    *axis = make_float4(cone.x / alen, cone.y / alen, cone.z / alen, cos_t);
    return true;
}

// The cell part, from the light part's axis = (unit axis, cos) and sin.
__device__ __forceinline__ bool hl_x_cone_cell(float4 c, float4 axis, float sin_t, float pad, float3 lo,
                                               float3 hi) {
#pragma clang fp contract(off)
    float mx = (lo.x + hi.x) * 0.5f, my = (lo.y + hi.y) * 0.5f, mz = (lo.z + hi.z) * 0.5f;
    float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
    float R = sqrtf(hl_x_dot3(ex, ey, ez, ex, ey, ez)) * 0.5f + pad;
    float vx = mx - c.x, vy = my - c.y, vz = mz - c.z;
    float vlen2 = hl_x_dot3(vx, vy, vz, vx, vy, vz);
    float v1 = hl_x_dot3(vx, vy, vz, axis.x, axis.y, axis.z);
    float perp = sqrtf(fmaxf(0.0f, vlen2 - v1 * v1));
    float dcp = axis.w * perp - v1 * sin_t;
    if (dcp > R) return false;
    if (v1 < -R) return false;
    return true;
}

// WGSL smoothstep, exact.
__device__ __forceinline__ float hl_x_smoothstep(float e0, float e1, float x) {
#pragma clang fp contract(off)
    float t = fminf(fmaxf((x - e0) / (e1 - e0), 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// max channel of radiance exactly as pbr_direct_light (deferred_lighting.wgsl:724-775)
// computes it for a point/spot light, and whether dist <= range.
__device__ __forceinline__ float hl_x_radiance_max(const HotLight& L, float wx, float wy, float wz,
                                                   bool* in_range) {
#pragma clang fp contract(off)
    float tx = L.p0.x - wx, ty = L.p0.y - wy, tz = L.p0.z - wz;
    float dist = sqrtf(hl_x_dot3(tx, ty, tz, tx, ty, tz));
    *in_range = !(dist > L.p0.w);
    float lx = tx / dist, ly = ty / dist, lz = tz / dist;
    float atten = 1.0f / (dist * dist + 0.0001f);
    float nd = dist / L.p0.w;
    atten = atten * fmaxf(0.0f, 1.0f - nd * nd * nd * nd);
    if (L.p2.w != kNoCone) {
        float cos_a = hl_x_dot3(-lx, -ly, -lz, L.p2.x, L.p2.y, L.p2.z);
        atten = atten * hl_x_smoothstep(L.p2.w, L.p1.w, cos_a);
    }
    float r = L.p1.x * atten, g = L.p1.y * atten, b = L.p1.z * atten;
    return fmaxf(fmaxf(r, g), b);
}

// a / b correctly rounded (the IEEE quotient) from y1, a reciprocal of b
// within one ulp refined once (hl_rcp_refined) or the correctly rounded
// reciprocal: two fma corrections of a * y1, the steps of the IEEE division
// expansion without its operand scaling and special-case fix-up.  Valid when
// no scaling or fix-up applies: a, b and a / b finite and normal, with the
// exponents of a and b less than 96 apart (hl_div_r_ok is a sufficient test).
// The same function of the same inputs as IEEE division, checked exhaustively
// for a = x + 0.5, b = W (every x < W <= 16384) and on 5e7 random operands.
__device__ __forceinline__ float hl_x_div_r(float a, float b, float y1) {
#pragma clang fp contract(off)
    const float q = a * y1;
    const float r = __builtin_fmaf(-b, q, a);
    const float q1 = __builtin_fmaf(r, y1, q);
    const float r1 = __builtin_fmaf(-b, q1, a);
    return __builtin_fmaf(r1, y1, q1);
}
__device__ __forceinline__ float hl_rcp_refined(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
}
// |b| in [2^-40, 2^40], every |a| in [2^-60, 2^40] (zero and NaN fail): the
// quotients are normal and the exponent gap is below 96.
__device__ __forceinline__ bool hl_div_r_ok(float b, float a0, float a1, float a2) {
    const float ab = fabsf(b);
    const float amax = fmaxf(fmaxf(fabsf(a0), fabsf(a1)), fabsf(a2));
    const float amin = fminf(fminf(fabsf(a0), fabsf(a1)), fabsf(a2));
    return ab >= 0x1p-40f && ab <= 0x1p40f && amax <= 0x1p40f && amin >= 0x1p-60f;
}

// hl_x_normalize with the three quotients from one refined reciprocal
// (hl_x_div_r) where that is exact: length in [2^-40, 2^40] and each
// component zero or at least 2^-60 in magnitude (a component's magnitude never
// exceeds the length).  A -0 component comes out +0; otherwise the bits of
// hl_x_normalize.  len (optional): the length.
__device__ __forceinline__ float3 hl_x_normalize_r(float x, float y, float z, float* len = nullptr) {
#pragma clang fp contract(off)
    const float l = sqrtf(hl_x_dot3(x, y, z, x, y, z));
    if (len) *len = l;
    const float m = fminf(fminf(x == 0.0f ? 1.0f : fabsf(x), y == 0.0f ? 1.0f : fabsf(y)),
                          z == 0.0f ? 1.0f : fabsf(z));
    if (l >= 0x1p-40f && l <= 0x1p40f && m >= 0x1p-60f) {
        const float y1 = hl_rcp_refined(l);
        return make_float3(hl_x_div_r(x, l, y1), hl_x_div_r(y, l, y1), hl_x_div_r(z, l, y1));
    }
    return make_float3(x / l, y / l, z / l);
}

// -------------------------------------------------------------- fast math
__device__ __forceinline__ float hl_f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hl_f_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float hl_f_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// floor(x / c) for x < 2^21, c >= 1, in four VALU operations instead of an
// integer division: (x + 0.5) / c is at least 0.5 / c away from an integer,
// and the estimate (reciprocal within 1 ulp, one fma rounding) is within
// 1.8e-7 relative, i.e. within (x + 0.5) * 1.8e-7 / c < 0.5 / c.
__device__ __forceinline__ uint32_t hl_floor_div(uint32_t x, uint32_t c) {
    const float ic = hl_f_rcp((float)c);
    return (uint32_t)__builtin_fmaf((float)x, ic, 0.5f * ic);
}

// ----------------------------------------------------------- texel decode
__device__ __forceinline__ float hl_half(uint16_t h) {
    _Float16 v;
    __builtin_memcpy(&v, &h, 2);
    return (float)v;
}

// Encoders of the G-buffer formats (the inverse direction of the decoders
// below): unorm8 = round-to-nearest-even of clamp(v, 0, 1) * 255, NaN -> 0;
// fp16 = IEEE round-to-nearest-even (v_cvt_f16_f32 in the default mode).
__device__ __forceinline__ uint32_t hl_enc_unorm8(float v) {
    const float c = fminf(fmaxf(v, 0.0f), 1.0f);  // maxNum: NaN -> 0
    return (uint32_t)__builtin_rintf(c * 255.0f);
}
__device__ __forceinline__ uint32_t hl_enc_half(float v) {
    const _Float16 h = (_Float16)v;
    uint16_t u;
    __builtin_memcpy(&u, &h, 2);
    return u;
}

// v / 255 exactly: v * (1/255) with the reciprocal split into its float part
// and the float of the remainder, fma(v, hi, v * lo); equal to the IEEE
// quotient for every v in 0..255 (checked in exact rational arithmetic).
__device__ __forceinline__ float hl_unorm8(uint32_t v) {
    const float hi = 0.003921568859368563f;     // float(1/255)
    const float lo = -2.319175823606301e-10f;   // float(1/255 - hi)
    const float fv = (float)v;
    return __builtin_fmaf(fv, hi, fv * lo);
}

typedef unsigned int hl_u32x2 __attribute__((ext_vector_type(2)));
typedef float hl_f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int hl_u32x4 __attribute__((ext_vector_type(4)));

// Global-memory view of a device pointer.  Accesses through it compile to
// global instructions even when the pointer itself was read from memory (a
// batched view's record), where the compiler would otherwise emit flat ones,
// which also count against the LDS/scalar wait counter.  Device memory only.
template <class T>
using hl_gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ hl_gptr<T> hl_g(T* p) {
    return (hl_gptr<T>)p;
}

__device__ __forceinline__ void hl_store_nt(float4* p, float4 v) {
    hl_f32x4 w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, hl_g(reinterpret_cast<hl_f32x4*>(p)));
}

// One output pixel in the target's format: RGBA32F, or RGBA16F with each
// channel rounded to the nearest binary16 (ties to even, overflow to inf).
__device__ __forceinline__ void hl_store_px(uint8_t* px, bool f16, float4 v) {
    if (f16) {
        const _Float16 h0 = (_Float16)v.x, h1 = (_Float16)v.y, h2 = (_Float16)v.z, h3 = (_Float16)v.w;
        typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
        u32x2 u;
        u.x = (uint32_t)__builtin_bit_cast(uint16_t, h0) | ((uint32_t)__builtin_bit_cast(uint16_t, h1) << 16);
        u.y = (uint32_t)__builtin_bit_cast(uint16_t, h2) | ((uint32_t)__builtin_bit_cast(uint16_t, h3) << 16);
        __builtin_nontemporal_store(u, hl_g(reinterpret_cast<u32x2*>(px)));
    } else {
        hl_store_nt(reinterpret_cast<float4*>(px), v);
    }
}

__device__ __forceinline__ float4 hl_ld4(const float4* p) {
    const hl_f32x4 v = *hl_g(reinterpret_cast<const hl_f32x4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}

// Texel fetch and decode are separate so that a caller can issue the loads of
// every G-buffer texture before the first one is waited for.
template <int FMT>
struct TexRaw {
    using type = hl_f32x4;
};
template <>
struct TexRaw<HL_FMT_RGBA8_UNORM> {
    using type = uint32_t;
};
template <>
struct TexRaw<HL_FMT_RGBA16_FLOAT> {
    using type = hl_u32x2;
};

// Texel (x, y) of t at a 32-bit byte offset from the base (every texture is
// below 4 GiB, checked on the host), so that the load takes the base from
// scalar registers and one 32-bit offset per lane.
template <int FMT>
__device__ __forceinline__ typename TexRaw<FMT>::type hl_fetch(const Tex& t, uint32_t x, uint32_t y) {
    using R = typename TexRaw<FMT>::type;
    return __builtin_nontemporal_load(hl_g((const R*)(t.data + (y * t.pitch + x * (uint32_t)sizeof(R)))));
}

template <int FMT>
__device__ __forceinline__ float4 hl_decode(typename TexRaw<FMT>::type v) {
    if constexpr (FMT == HL_FMT_RGBA8_UNORM) {
        return make_float4(hl_unorm8(v & 0xffu), hl_unorm8((v >> 8) & 0xffu),
                           hl_unorm8((v >> 16) & 0xffu), hl_unorm8(v >> 24));
    } else if constexpr (FMT == HL_FMT_RGBA16_FLOAT) {
        return make_float4(hl_half((uint16_t)(v.x & 0xffffu)), hl_half((uint16_t)(v.x >> 16)),
                           hl_half((uint16_t)(v.y & 0xffffu)), hl_half((uint16_t)(v.y >> 16)));
    } else {
        return make_float4(v.x, v.y, v.z, v.w);
    }
}

template <int FMT>
__device__ __forceinline__ float4 hl_load_rgba(const Tex& t, uint32_t x, uint32_t y) {
    return hl_decode<FMT>(hl_fetch<FMT>(t, x, y));
}

__device__ __forceinline__ float hl_load_depth(const Tex& t, uint32_t x, uint32_t y) {
    return __builtin_nontemporal_load(hl_g((const float*)(t.data + (y * t.pitch + x * 4u))));
}

// -------------------------------------------------------------- wave ops
__device__ __forceinline__ uint64_t hl_ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t hl_lane() { return __lane_id(); }

// number of set bits of m below this lane
__device__ __forceinline__ uint32_t hl_prefix(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <typename T>
__device__ __forceinline__ T hl_readlane(T v, uint32_t lane) {
    static_assert(sizeof(T) == 4, "32-bit only");
    int iv;
    __builtin_memcpy(&iv, &v, 4);
    iv = __builtin_amdgcn_readlane(iv, (int)lane);
    T r;
    __builtin_memcpy(&r, &iv, 4);
    return r;
}

template <typename T>
__device__ __forceinline__ T hl_uniform(T v) {
    static_assert(sizeof(T) == 4, "32-bit only");
    int iv;
    __builtin_memcpy(&iv, &v, 4);
    iv = __builtin_amdgcn_readfirstlane(iv);
    T r;
    __builtin_memcpy(&r, &iv, 4);
    return r;
}

__device__ __forceinline__ float hl_wave_min(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, HL_WAVE));
    return v;
}
__device__ __forceinline__ float hl_wave_max(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, HL_WAVE));
    return v;
}
__device__ __forceinline__ float hl_wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, HL_WAVE);
    return v;
}

// ---------------------------------------------------- DPP wave reductions
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float hl_dpp(float old, float src) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src),
                                                       CTRL, ROW_MASK, 0xF, false));
}

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int hl_dpp_i(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, CTRL, ROW_MASK, 0xF, false);
}

// Order-preserving map of non-NaN floats onto signed integers (and back: the
// map is an involution): positive floats keep their bits, negative floats
// flip their magnitude bits.  -0 maps below +0.
__device__ __forceinline__ int hl_fkey(float x) {
    const int b = __float_as_int(x);
    return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float hl_funkey(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }

// Full-wave reductions: quad swaps, half-row and row mirrors, then the
// row_bcast:15 / row_bcast:31 steps; lane 63 holds the result, which is read
// back as a wave-uniform value.  All 64 lanes must be active.  Integer min/max
// with their identity as the DPP "old" operand fold each step into one
// v_min_i32_dpp / v_max_i32_dpp (float min/max would add a move, a NaN
// quieting max and hazard nops per step).
__device__ __forceinline__ int dpp_min_key(int v) {
    constexpr int I = 0x7fffffff;
    v = min(v, hl_dpp_i<0xB1>(I, v));
    v = min(v, hl_dpp_i<0x4E>(I, v));
    v = min(v, hl_dpp_i<0x141>(I, v));
    v = min(v, hl_dpp_i<0x140>(I, v));
    v = min(v, hl_dpp_i<0x142, 0xA>(I, v));
    v = min(v, hl_dpp_i<0x143, 0xC>(I, v));
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int dpp_max_key(int v) {
    constexpr int I = (int)0x80000000;
    v = max(v, hl_dpp_i<0xB1>(I, v));
    v = max(v, hl_dpp_i<0x4E>(I, v));
    v = max(v, hl_dpp_i<0x141>(I, v));
    v = max(v, hl_dpp_i<0x140>(I, v));
    v = max(v, hl_dpp_i<0x142, 0xA>(I, v));
    v = max(v, hl_dpp_i<0x143, 0xC>(I, v));
    return __builtin_amdgcn_readlane(v, 63);
}
// min / max over the lanes with `take` set (+inf / -inf when none); inputs
// must not be NaN.
__device__ __forceinline__ float dpp_min_if(bool take, float v) {
    return hl_funkey(dpp_min_key(take ? hl_fkey(v) : 0x7fffffff));
}
__device__ __forceinline__ float dpp_max_if(bool take, float v) {
    return hl_funkey(dpp_max_key(take ? hl_fkey(v) : (int)0x80000000));
}
__device__ __forceinline__ float dpp_min(float v) { return dpp_min_if(true, v); }
__device__ __forceinline__ float dpp_max(float v) { return dpp_max_if(true, v); }
__device__ __forceinline__ float dpp_sum(float v) {
    v = v + hl_dpp<0xB1>(0.0f, v);
    v = v + hl_dpp<0x4E>(0.0f, v);
    v = v + hl_dpp<0x141>(0.0f, v);
    v = v + hl_dpp<0x140>(0.0f, v);
    v = v + hl_dpp<0x142, 0xA>(0.0f, v);
    v = v + hl_dpp<0x143, 0xC>(0.0f, v);
    return hl_readlane(v, 63);
}

// Placed in a rarely taken block, keeps the optimizer from executing the block
// unconditionally and selecting its result (an empty volatile asm statement
// has side effects, so the block cannot be speculated).
__device__ __forceinline__ void hl_rare() { __asm__ volatile(""); }

// Wave-local ordering point between LDS writes and reads of other lanes.
__device__ __forceinline__ void hl_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// blockIdx -> work index so consecutive work lands on one XCD (blocks are
// dispatched round-robin over the 8 XCDs of an MI355X).
__device__ __forceinline__ uint32_t hl_xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t nx = 8;
    uint32_t q = nb / nx, r = nb % nx;
    uint32_t xcd = b % nx, k = b / nx;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

}  // namespace hl
