// hl_cull.hip -- light preparation and light assignment kernels for gfx950.
//
//  k_prep_lights    GpuLight[] -> HotLight[] (shading) + CullLight[] (culling)
//  k_cull_tiles     reference 16-px tile lists (light_cull.wgsl:123-180), one
//                   wavefront per tile, ballot + prefix-popcount compaction
//  k_cluster_*      log-depth froxel lists: per-column pre-cull and per-slice
//                   counts, one exclusive scan, then the fill pass
#include "hl_internal.hpp"

namespace hl {

// ------------------------------------------------------------------- prep
__global__ __launch_bounds__(256) void k_prep_lights(const GpuLight* __restrict__ lights,
                                                     const uint32_t* __restrict__ subset,
                                                     uint32_t n,
                                                     const GpuCameraUniforms* __restrict__ cam,
                                                     HotLight* __restrict__ hot,
                                                     CullLight* __restrict__ cull,
                                                     uint32_t* __restrict__ flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GpuLight* l = &lights[subset ? subset[i] : i];
    const float4 pr = *reinterpret_cast<const float4*>(l->position_range);
    const float4 dro = *reinterpret_cast<const float4*>(l->direction_outer);
    const float4 ci = *reinterpret_cast<const float4*>(l->color_intensity);
    const uint32_t type = l->light_type;
    const float inner = l->inner_angle;
    const float inf = __builtin_inff();
    HotLight h;
    float range = type == HL_LIGHT_DIRECTIONAL ? inf : pr.w;
    h.p0 = make_float4(pr.x, pr.y, pr.z, range);
    {
#pragma clang fp contract(off)
        h.p1 = make_float4(ci.x * ci.w, ci.y * ci.w, ci.z * ci.w, inner);
    }
    h.p2 = make_float4(dro.x, dro.y, dro.z, type == HL_LIGHT_SPOT ? dro.w : kNoCone);
    h.p3 = ci;
    hot[i] = h;
    if (cull) {
        const float* v = cam->view;
        CullLight c;
        c.vpos = make_float4(hl_x_mat_row(v, 0, pr.x, pr.y, pr.z, 1.0f),
                             hl_x_mat_row(v, 1, pr.x, pr.y, pr.z, 1.0f),
                             hl_x_mat_row(v, 2, pr.x, pr.y, pr.z, 1.0f), range);
        c.cone = make_float4(hl_x_mat_row(v, 0, dro.x, dro.y, dro.z, 0.0f),
                             hl_x_mat_row(v, 1, dro.x, dro.y, dro.z, 0.0f),
                             hl_x_mat_row(v, 2, dro.x, dro.y, dro.z, 0.0f),
                             type == HL_LIGHT_SPOT ? dro.w : kNoCone);
        cull[i] = c;
    }
    if (flags) {
        uint32_t f = 0;
        // shadow_factor samples the atlas only below MAX_SHADOW_LIGHTS = 42
        // (deferred_lighting.wgsl:453-459)
        if (l->shadow_index != HL_NO_SHADOW && i < 42u) f |= kFlagShadowIgnored;
        if (l->ies_profile_index >= 0 || l->light_function_index >= 0) f |= kFlagTexturedIgnored;
        if (f) atomicOr(flags, f);
    }
}

// -------------------------------------------------------------- tile cull
// One wavefront per 16x16 tile; four tiles per 256-thread block.
__global__ __launch_bounds__(256) void k_cull_tiles(const GpuCameraUniforms* __restrict__ cam,
                                                    LightCullParams p,
                                                    const CullLight* __restrict__ cl,
                                                    uint32_t* __restrict__ lists,
                                                    uint32_t* __restrict__ counts) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t total = p.num_tiles_x * p.num_tiles_y;
    if (tile >= total) return;
    const uint32_t tile_x = tile % p.num_tiles_x, tile_y = tile / p.num_tiles_x;
    float3 pl[4];
    {
#pragma clang fp contract(off)
        const float fw = (float)p.screen_width, fh = (float)p.screen_height;
        const float ts = (float)HL_TILE_SIZE;
        float ndc_left = ((float)tile_x * ts / fw) * 2.0f - 1.0f;
        float ndc_right = fminf(((float)(tile_x + 1u) * ts / fw) * 2.0f - 1.0f, 1.0f);
        float ndc_top = 1.0f - ((float)tile_y * ts / fh) * 2.0f;
        float ndc_bottom = fmaxf(1.0f - ((float)(tile_y + 1u) * ts / fh) * 2.0f, -1.0f);
        const float p00 = cam->proj[0], p11 = cam->proj[5];
        pl[0] = hl_x_plane(p00, p11, ndc_left, ndc_bottom, ndc_left, ndc_top);
        pl[1] = hl_x_plane(p00, p11, ndc_right, ndc_top, ndc_right, ndc_bottom);
        pl[2] = hl_x_plane(p00, p11, ndc_left, ndc_top, ndc_right, ndc_top);
        pl[3] = hl_x_plane(p00, p11, ndc_right, ndc_bottom, ndc_left, ndc_bottom);
    }
    uint32_t count = 0;
    uint32_t* out = lists + (size_t)tile * HL_MAX_LIGHTS_PER_TILE;
    for (uint32_t base = 0; base < p.num_lights; base += 64u) {
        const uint32_t i = base + lane;
        bool pass = false;
        if (i < p.num_lights) {
            const float4 c = cl[i].vpos;
            pass = hl_x_plane_pass(pl[0], c, c.w) && hl_x_plane_pass(pl[1], c, c.w) &&
                   hl_x_plane_pass(pl[2], c, c.w) && hl_x_plane_pass(pl[3], c, c.w);
        }
        const uint64_t m = hl_ballot(pass);
        const uint32_t pos = count + hl_prefix(m);
        if (pass && pos < HL_MAX_LIGHTS_PER_TILE) out[pos] = i;
        count += (uint32_t)__popcll(m);
        if (count >= HL_MAX_LIGHTS_PER_TILE) break;
    }
    if (lane == 0) counts[tile] = count < HL_MAX_LIGHTS_PER_TILE ? count : HL_MAX_LIGHTS_PER_TILE;
}

// ------------------------------------------------------------ clusters
struct ColumnGeom {
    float3 pl[4];
    float l, r, b, t;
};

// The column's screen rectangle in NDC and its lateral planes (same
// construction as the tile cull).  false when the column lies off screen.
__device__ __forceinline__ bool column_geom(const GpuCameraUniforms* cam, const ClusterArgs& a,
                                            uint32_t cx, uint32_t cy, ColumnGeom* g) {
#pragma clang fp contract(off)
    uint32_t px0 = cx * a.cell_w, py0 = cy * a.cell_h;
    if (px0 >= a.width || py0 >= a.height) return false;
    uint32_t px1 = min(px0 + a.cell_w, a.width), py1 = min(py0 + a.cell_h, a.height);
    const float fw = (float)a.width, fh = (float)a.height;
    g->l = ((float)px0 / fw) * 2.0f - 1.0f;
    g->r = ((float)px1 / fw) * 2.0f - 1.0f;
    g->t = 1.0f - ((float)py0 / fh) * 2.0f;
    g->b = 1.0f - ((float)py1 / fh) * 2.0f;
    const float p00 = cam->proj[0], p11 = cam->proj[5];
    g->pl[0] = hl_x_plane(p00, p11, g->l, g->b, g->l, g->t);
    g->pl[1] = hl_x_plane(p00, p11, g->r, g->t, g->r, g->b);
    g->pl[2] = hl_x_plane(p00, p11, g->l, g->t, g->r, g->t);
    g->pl[3] = hl_x_plane(p00, p11, g->r, g->b, g->l, g->b);
    return true;
}

__device__ __forceinline__ void slice_aabb(const GpuCameraUniforms* cam, const ColumnGeom& g,
                                           float zn, float zf, float3* lo, float3* hi) {
#pragma clang fp contract(off)
    const float p00 = cam->proj[0], p11 = cam->proj[5];
    float x0 = g.l * zn / p00, x1 = g.r * zn / p00, x2 = g.l * zf / p00, x3 = g.r * zf / p00;
    float y0 = g.b * zn / p11, y1 = g.t * zn / p11, y2 = g.b * zf / p11, y3 = g.t * zf / p11;
    lo->x = fminf(fminf(fminf(x0, x1), x2), x3);
    hi->x = fmaxf(fmaxf(fmaxf(x0, x1), x2), x3);
    lo->y = fminf(fminf(fminf(y0, y1), y2), y3);
    hi->y = fmaxf(fmaxf(fmaxf(y0, y1), y2), y3);
    lo->z = -zf;
    hi->z = -zn;
}

__device__ __forceinline__ bool slice_member(const CullLight& c, float pad, float3 lo, float3 hi) {
#pragma clang fp contract(off)
    if (!hl_x_sphere_aabb(c.vpos, c.vpos.w + pad, lo, hi)) return false;
    if (c.cone.w != kNoCone && !hl_x_cone_aabb(c.vpos, c.cone, pad, lo, hi)) return false;
    return true;
}

// Block per column: build the column list (lateral-plane pre-cull, ascending)
// then count the members of every slice.
__global__ __launch_bounds__(256) void k_cluster_columns(ClusterArgs a,
                                                         const GpuCameraUniforms* __restrict__ cam,
                                                         const CullLight* __restrict__ cl,
                                                         uint32_t* __restrict__ col_list,
                                                         uint32_t* __restrict__ col_count,
                                                         uint32_t* __restrict__ raw) {
    __shared__ uint32_t s_cnt[4];
    const uint32_t col = blockIdx.x;
    const uint32_t ncol = a.grid_x * a.grid_y;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    ColumnGeom g;
    const bool on = column_geom(cam, a, col % a.grid_x, col / a.grid_x, &g);
    uint32_t* list = col_list + (size_t)col * a.num_lights;
    uint32_t total = 0;
    if (on) {
        for (uint32_t base = 0; base < a.num_lights; base += 256u) {
            const uint32_t i = base + threadIdx.x;
            bool pass = false;
            if (i < a.num_lights) {
                const float4 c = cl[i].vpos;
                const float r = c.w + a.r_pad_column;
                pass = hl_x_plane_pass(g.pl[0], c, r) && hl_x_plane_pass(g.pl[1], c, r) &&
                       hl_x_plane_pass(g.pl[2], c, r) && hl_x_plane_pass(g.pl[3], c, r);
            }
            const uint64_t m = hl_ballot(pass);
            if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t off = total;
            for (uint32_t w = 0; w < wave; ++w) off += s_cnt[w];
            if (pass) list[off + hl_prefix(m)] = i;
            total += s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) col_count[col] = total;
    __threadfence_block();
    __syncthreads();
    for (uint32_t k = wave; k < a.grid_z; k += 4u) {
        uint32_t hits = 0;
        if (on) {
            float3 lo, hi;
            slice_aabb(cam, g, a.slice_near[k], a.slice_far[k], &lo, &hi);
            const float pad = a.slice_r_pad[k];
            for (uint32_t j = 0; j < total; j += 64u) {
                bool mem = false;
                if (j + lane < total) mem = slice_member(cl[list[j + lane]], pad, lo, hi);
                hits += (uint32_t)__popcll(hl_ballot(mem));
            }
        }
        if (lane == 0) raw[col + ncol * k] = hits;
    }
}

// Single block: capped counts -> offsets (exclusive scan), pool truncation, flags.
__global__ __launch_bounds__(1024) void k_cluster_scan(const uint32_t* __restrict__ raw,
                                                       uint32_t ncell, uint32_t cap,
                                                       uint32_t pool_cap,
                                                       uint32_t* __restrict__ offsets,
                                                       uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ flags_out) {
    __shared__ unsigned long long s_sum[1024];
    __shared__ uint32_t s_flags;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (ncell + 1023u) / 1024u;
    const uint32_t c0 = t * per, c1 = min(c0 + per, ncell);
    if (t == 0) s_flags = 0;
    unsigned long long local = 0;
    uint32_t fl = 0;
    for (uint32_t c = c0; c < c1; ++c) {
        uint32_t n = raw[c];
        if (cap && n > cap) {
            n = cap;
            fl |= HL_OVERFLOW_CAP;
        }
        local += n;
    }
    s_sum[t] = local;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partial sums
    for (uint32_t o = 1; o < 1024u; o <<= 1) {
        unsigned long long v = t >= o ? s_sum[t - o] : 0ull;
        __syncthreads();
        s_sum[t] += v;
        __syncthreads();
    }
    unsigned long long off = s_sum[t] - local;
    for (uint32_t c = c0; c < c1; ++c) {
        uint32_t n = raw[c];
        if (cap && n > cap) n = cap;
        uint32_t o = off > pool_cap ? pool_cap : (uint32_t)off;
        uint32_t room = pool_cap - o;
        uint32_t nn = n;
        if (nn > room) {
            nn = room;
            fl |= HL_OVERFLOW_POOL;
        }
        offsets[c] = o;
        counts[c] = nn;
        off += n;
    }
    if (fl) atomicOr(&s_flags, fl);
    __syncthreads();
    if (t == 0) flags_out[0] |= s_flags;
}

// Block per column, wave per slice: write the members of each slice, ascending,
// stopping at the (capped, pool-truncated) count.
__global__ __launch_bounds__(256) void k_cluster_fill(ClusterArgs a,
                                                      const GpuCameraUniforms* __restrict__ cam,
                                                      const CullLight* __restrict__ cl,
                                                      const uint32_t* __restrict__ col_list,
                                                      const uint32_t* __restrict__ col_count,
                                                      const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ pool) {
    const uint32_t col = blockIdx.x;
    const uint32_t ncol = a.grid_x * a.grid_y;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    ColumnGeom g;
    if (!column_geom(cam, a, col % a.grid_x, col / a.grid_x, &g)) return;
    const uint32_t* list = col_list + (size_t)col * a.num_lights;
    const uint32_t total = col_count[col];
    for (uint32_t k = wave; k < a.grid_z; k += 4u) {
        const uint32_t c = col + ncol * k;
        const uint32_t cnt = counts[c];
        if (cnt == 0) continue;
        uint32_t* dst = pool + offsets[c];
        float3 lo, hi;
        slice_aabb(cam, g, a.slice_near[k], a.slice_far[k], &lo, &hi);
        const float pad = a.slice_r_pad[k];
        uint32_t written = 0;
        for (uint32_t j = 0; j < total && written < cnt; j += 64u) {
            bool mem = false;
            uint32_t idx = 0;
            if (j + lane < total) {
                idx = list[j + lane];
                mem = slice_member(cl[idx], pad, lo, hi);
            }
            const uint64_t m = hl_ballot(mem);
            const uint32_t pos = written + hl_prefix(m);
            if (mem && pos < cnt) dst[pos] = idx;
            written += (uint32_t)__popcll(m);
        }
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------- launchers
hipError_t launch_prep(hipStream_t s, const GpuLight* lights, const uint32_t* subset, uint32_t n,
                       const GpuCameraUniforms* cam, HotLight* hot, CullLight* cull,
                       uint32_t* flags) {
    if (n == 0) return hipSuccess;
    k_prep_lights<<<(n + 255) / 256, 256, 0, s>>>(lights, subset, n, cam, hot, cull, flags);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    k_fill_u32<<<(unsigned)blocks, 256, 0, s>>>(p, v, n);
    return hipGetLastError();
}

hipError_t launch_cull_tiles(hipStream_t s, const GpuCameraUniforms* cam, const LightCullParams& p,
                             const CullLight* cl, uint32_t* lists, uint32_t* counts) {
    const uint32_t total = p.num_tiles_x * p.num_tiles_y;
    if (total == 0) return hipSuccess;
    k_cull_tiles<<<(total + 3) / 4, 256, 0, s>>>(cam, p, cl, lists, counts);
    return hipGetLastError();
}

hipError_t launch_cluster_cull(hipStream_t s, const ClusterArgs& a, const GpuCameraUniforms* cam,
                               const CullLight* cl, uint32_t* col_list, uint32_t* col_count,
                               uint32_t* raw, uint32_t* offsets, uint32_t* counts, uint32_t* pool,
                               uint32_t pool_cap, uint32_t* flags) {
    const uint32_t ncol = a.grid_x * a.grid_y;
    const uint32_t ncell = ncol * a.grid_z;
    k_cluster_columns<<<ncol, 256, 0, s>>>(a, cam, cl, col_list, col_count, raw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_cluster_scan<<<1, 1024, 0, s>>>(raw, ncell, a.cap, pool_cap, offsets, counts, flags);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_cluster_fill<<<ncol, 256, 0, s>>>(a, cam, cl, col_list, col_count, offsets, counts, pool);
    return hipGetLastError();
}

}  // namespace hl
