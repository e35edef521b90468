// hl_gbuffer.hip -- the G-buffer producer's format contract for gfx950: packs
// per-pixel surface attributes into the textures the deferred pass reads,
// with the specular F0 resolved into their alpha channels, as the G-buffer
// pass's fragment shader writes them (crates/passes/3d/helio-pass-gbuffer/
// shaders/gbuffer.wgsl:10-15, 308-329, 480-493).
//
//  k_gbuffer_pack   one thread per pixel, a block per 64 x 4 pixels: four
//                   (five with a specular F0) 16-B loads, four stores of the
//                   target formats, and the lightmap sentinel when asked.
//                   HBM-bound: 64-80 B read and 24 B (reference formats) to
//                   64 B (fp32) written per pixel.
#include "hl_internal.hpp"

namespace hl {

namespace {

__device__ __forceinline__ float4 load_plane(const float* p, uint32_t pitch, uint32_t x, uint32_t y) {
    return hl_ld4(reinterpret_cast<const float4*>(reinterpret_cast<const uint8_t*>(p) + (size_t)y * pitch) + x);
}

// One texel of format FMT at (x, y) of t.
template <int FMT>
__device__ __forceinline__ void store_texel(const Tex& t, uint32_t x, uint32_t y, float4 v) {
    uint8_t* row = const_cast<uint8_t*>(t.data) + (size_t)y * t.pitch;
    if constexpr (FMT == HL_FMT_RGBA8_UNORM) {
        const uint32_t w = hl_enc_unorm8(v.x) | (hl_enc_unorm8(v.y) << 8) | (hl_enc_unorm8(v.z) << 16) |
                           (hl_enc_unorm8(v.w) << 24);
        *hl_g(reinterpret_cast<uint32_t*>(row) + x) = w;
    } else if constexpr (FMT == HL_FMT_RGBA16_FLOAT) {
        hl_u32x2 w;
        w.x = hl_enc_half(v.x) | (hl_enc_half(v.y) << 16);
        w.y = hl_enc_half(v.z) | (hl_enc_half(v.w) << 16);
        *hl_g(reinterpret_cast<hl_u32x2*>(row) + x) = w;
    } else {
        hl_f32x4 w = {v.x, v.y, v.z, v.w};
        *hl_g(reinterpret_cast<hl_f32x4*>(row) + x) = w;
    }
}

struct PackArgs {
    uint32_t width, height, src_pitch;
    const float *albedo, *normal, *orm, *emissive, *f0;
    Tex out_albedo, out_normal, out_orm, out_emissive, out_lightmap;
};

// gbuffer.wgsl:324-328: clamp(mix(vec3(0.04), albedo, metallic), 0, 0.999),
// mix(a, b, t) = a * (1 - t) + b * t, one IEEE operation at a time.
__device__ __forceinline__ float f0_metallic(float albedo, float metallic) {
#pragma clang fp contract(off)
    const float m = 0.04f * (1.0f - metallic) + albedo * metallic;
    return fminf(fmaxf(m, 0.0f), 0.999f);
}

template <int FA, int FN>
__global__ __launch_bounds__(256) void k_gbuffer_pack(PackArgs a) {
    const uint32_t x = blockIdx.x * 64u + (threadIdx.x & 63u);
    const uint32_t y = blockIdx.y * 4u + (threadIdx.x >> 6);
    if (x >= a.width || y >= a.height) return;
    const float4 alb = load_plane(a.albedo, a.src_pitch, x, y);
    const float4 n = load_plane(a.normal, a.src_pitch, x, y);
    const float4 orm = load_plane(a.orm, a.src_pitch, x, y);
    const float4 em = load_plane(a.emissive, a.src_pitch, x, y);
    float f0r, f0g, f0b;
    if (a.f0) {
        const float4 f = load_plane(a.f0, a.src_pitch, x, y);
        f0r = f.x;
        f0g = f.y;
        f0b = f.z;
    } else {
        f0r = f0_metallic(alb.x, orm.z);
        f0g = f0_metallic(alb.y, orm.z);
        f0b = f0_metallic(alb.z, orm.z);
    }
    store_texel<FA>(a.out_albedo, x, y, alb);
    store_texel<FN>(a.out_normal, x, y, make_float4(n.x, n.y, n.z, f0r));
    store_texel<FA>(a.out_orm, x, y, make_float4(orm.x, orm.y, orm.z, f0g));
    store_texel<FN>(a.out_emissive, x, y, make_float4(em.x, em.y, em.z, f0b));
    if (a.out_lightmap.data) {
        const uint32_t s = hl_enc_half(-1.0f);
        *hl_g(reinterpret_cast<uint32_t*>(const_cast<uint8_t*>(a.out_lightmap.data) +
                                          (size_t)y * a.out_lightmap.pitch) + x) = s | (s << 16);
    }
}

template <int FA, int FN>
hipError_t launch_pack_fmt(hipStream_t s, const PackArgs& a) {
    const dim3 grid((a.width + 63u) / 64u, (a.height + 3u) / 4u);
    k_gbuffer_pack<FA, FN><<<grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_gbuffer_pack(hipStream_t s, int gbuf_set, const hl_surface& sf, const hl_gbuffer& gb) {
    PackArgs a;
    a.width = sf.width;
    a.height = sf.height;
    a.src_pitch = sf.pitch_bytes;
    a.albedo = sf.albedo_d;
    a.normal = sf.normal_d;
    a.orm = sf.orm_d;
    a.emissive = sf.emissive_d;
    a.f0 = sf.specular_f0_d;
    a.out_albedo = Tex{(const uint8_t*)gb.albedo.data_d, gb.albedo.pitch_bytes};
    a.out_normal = Tex{(const uint8_t*)gb.normal.data_d, gb.normal.pitch_bytes};
    a.out_orm = Tex{(const uint8_t*)gb.orm.data_d, gb.orm.pitch_bytes};
    a.out_emissive = Tex{(const uint8_t*)gb.emissive.data_d, gb.emissive.pitch_bytes};
    a.out_lightmap = Tex{(const uint8_t*)gb.lightmap_uv.data_d, gb.lightmap_uv.pitch_bytes};
    if (sf.width == 0 || sf.height == 0) return hipSuccess;
    switch (gbuf_set) {
        case kSetRef: return launch_pack_fmt<HL_FMT_RGBA8_UNORM, HL_FMT_RGBA16_FLOAT>(s, a);
        case kSet16: return launch_pack_fmt<HL_FMT_RGBA16_FLOAT, HL_FMT_RGBA16_FLOAT>(s, a);
        default: return launch_pack_fmt<HL_FMT_RGBA32_FLOAT, HL_FMT_RGBA32_FLOAT>(s, a);
    }
}

}  // namespace hl
