// hl_cull.hip -- light preparation and light assignment kernels for gfx950.
//
//  k_prep_lights      GpuLight[] -> HotLight[] (shading) + CullLight[] (culling)
//  k_cull_tiles       reference 16-px tile lists (light_cull.wgsl:123-180), one
//                     wavefront per tile, ballot + prefix-popcount compaction
//  k_cluster_geom     per column: NDC rectangle and lateral planes; per cell:
//                     the view-space AABB
//  k_cluster_mark     light-parallel: each light bounds the cells it can reach
//                     (projected bounding box x slice range); a wavefront tests
//                     those candidates 64 at a time with the exact membership
//                     rule and sets the member bits in a two-level bitmask
//                     (bits + a summary bit per non-empty word)
//  k_cluster_count    per cluster: popcount over its non-empty words
//  k_cluster_scan     exclusive scan of the capped counts (single block)
//  k_cluster_fill     per cluster: ascending indices from the bitmask, which it
//                     clears again for the next call
// Membership is decided by the same exact FP32 tests whatever the order of
// evaluation, so the lists equal the per-cluster brute force bit for bit.
#include "hl_internal.hpp"

namespace hl {

// ------------------------------------------------------------------- prep
__device__ __forceinline__ void prep_one(const GpuLight* __restrict__ lights,
                                         const uint32_t* __restrict__ subset, uint32_t n,
                                         const GpuCameraUniforms* __restrict__ cam,
                                         HotLight* __restrict__ hot, CullLight* __restrict__ cull,
                                         uint32_t* __restrict__ flags, uint32_t i) {
    if (i >= n) return;
    const GpuLight* l = &lights[subset ? subset[i] : i];
    const float4 pr = *reinterpret_cast<const float4*>(l->position_range);
    const float4 dro = *reinterpret_cast<const float4*>(l->direction_outer);
    const float4 ci = *reinterpret_cast<const float4*>(l->color_intensity);
    const uint32_t type = l->light_type;
    const float inner = l->inner_angle;
    const float inf = __builtin_inff();
    const float range = type == HL_LIGHT_DIRECTIONAL ? inf : pr.w;
    if (hot) {
        HotLight h;
        h.p0 = make_float4(pr.x, pr.y, pr.z, range);
        {
#pragma clang fp contract(off)
            h.p1 = make_float4(ci.x * ci.w, ci.y * ci.w, ci.z * ci.w, inner);
        }
        h.p2 = make_float4(dro.x, dro.y, dro.z, type == HL_LIGHT_SPOT ? dro.w : kNoCone);
        h.p3 = ci;
        hot[i] = h;
    }
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

__global__ __launch_bounds__(256) void k_prep_lights(const GpuLight* __restrict__ lights,
                                                     const uint32_t* __restrict__ subset,
                                                     uint32_t n,
                                                     const GpuCameraUniforms* __restrict__ cam,
                                                     HotLight* __restrict__ hot,
                                                     CullLight* __restrict__ cull,
                                                     uint32_t* __restrict__ flags) {
    prep_one(lights, subset, n, cam, hot, cull, flags, blockIdx.x * blockDim.x + threadIdx.x);
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
    uint32_t on;
    uint32_t pad[3];
};

struct CellBox {
    float lo[3], hi[3];
};

// geometry scratch: ColumnGeom per column, then (256-B aligned) CellBox per cell
__host__ __device__ inline size_t cluster_geom_offset_boxes_dev(uint32_t ncol) {
    return (sizeof(ColumnGeom) * (size_t)ncol + 255) & ~(size_t)255;
}

__device__ __forceinline__ bool column_rect(const ClusterArgs& a, uint32_t cx, uint32_t cy,
                                            float* l, float* r, float* b, float* t) {
#pragma clang fp contract(off)
    const uint32_t px0 = cx * a.cell_w, py0 = cy * a.cell_h;
    if (px0 >= a.width || py0 >= a.height) return false;
    const uint32_t px1 = min(px0 + a.cell_w, a.width), py1 = min(py0 + a.cell_h, a.height);
    const float fw = (float)a.width, fh = (float)a.height;
    *l = ((float)px0 / fw) * 2.0f - 1.0f;
    *r = ((float)px1 / fw) * 2.0f - 1.0f;
    *t = 1.0f - ((float)py0 / fh) * 2.0f;
    *b = 1.0f - ((float)py1 / fh) * 2.0f;
    return true;
}

// One thread per cell: the cell's view-space AABB (the column's NDC corners
// unprojected at the slice's padded near and far distances); the slice-0
// thread of each column also writes the column's lateral planes (the tile
// cull's construction on the column's pixel rectangle).
__device__ __forceinline__ void geom_cell(const ClusterArgs& a,
                                          const GpuCameraUniforms* __restrict__ cam,
                                          ColumnGeom* __restrict__ geom,
                                          CellBox* __restrict__ boxes, uint32_t cell) {
#pragma clang fp contract(off)
    const uint32_t ncol = a.grid_x * a.grid_y;
    if (cell >= ncol * a.grid_z) return;
    const uint32_t col = cell % ncol, k = cell / ncol;
    const uint32_t cx = col % a.grid_x, cy = col / a.grid_x;
    float l = 0.0f, r = 0.0f, b = 0.0f, t = 0.0f;
    const bool on = column_rect(a, cx, cy, &l, &r, &b, &t);
    const float p00 = cam->proj[0], p11 = cam->proj[5];
    CellBox bx;
    if (on) {
        const float zn = a.slice_near[k], zf = a.slice_far[k];
        const float x0 = l * zn / p00, x1 = r * zn / p00, x2 = l * zf / p00, x3 = r * zf / p00;
        const float y0 = b * zn / p11, y1 = t * zn / p11, y2 = b * zf / p11, y3 = t * zf / p11;
        bx.lo[0] = fminf(fminf(fminf(x0, x1), x2), x3);
        bx.hi[0] = fmaxf(fmaxf(fmaxf(x0, x1), x2), x3);
        bx.lo[1] = fminf(fminf(fminf(y0, y1), y2), y3);
        bx.hi[1] = fmaxf(fmaxf(fmaxf(y0, y1), y2), y3);
        bx.lo[2] = -zf;
        bx.hi[2] = -zn;
    } else {
        bx.lo[0] = bx.lo[1] = bx.lo[2] = bx.hi[0] = bx.hi[1] = bx.hi[2] = 0.0f;
    }
    boxes[cell] = bx;
    if (k != 0) return;
    ColumnGeom g;
    g.on = on ? 1u : 0u;
    g.pad[0] = g.pad[1] = g.pad[2] = 0;
    g.l = l;
    g.r = r;
    g.b = b;
    g.t = t;
    if (on) {
        g.pl[0] = hl_x_plane(p00, p11, l, b, l, t);
        g.pl[1] = hl_x_plane(p00, p11, r, t, r, b);
        g.pl[2] = hl_x_plane(p00, p11, l, t, r, t);
        g.pl[3] = hl_x_plane(p00, p11, r, b, l, b);
    } else {
        for (int i = 0; i < 4; ++i) g.pl[i] = make_float3(0.0f, 0.0f, 0.0f);
    }
    geom[col] = g;
}

// first k in [0, n) with v[k] >= x (v ascending); n if none
__device__ __forceinline__ int lower_bound(const float* v, int n, float x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Membership of light (c) in cell (col, k): the definition of the clustered
// lists -- lateral planes of the column with the column padding, the cell's
// AABB with the slice padding, and the cone for spot lights.
__device__ __forceinline__ bool cell_member(const ClusterArgs& a, const CullLight& c,
                                            const ColumnGeom& g, const CellBox& bx, uint32_t k) {
#pragma clang fp contract(off)
    const float pad = a.slice_r_pad[k];
    const float rc = c.vpos.w + pad;
    if (!(hl_x_plane_pass(g.pl[0], c.vpos, rc) && hl_x_plane_pass(g.pl[1], c.vpos, rc) &&
          hl_x_plane_pass(g.pl[2], c.vpos, rc) && hl_x_plane_pass(g.pl[3], c.vpos, rc)))
        return false;
    const float3 lo = make_float3(bx.lo[0], bx.lo[1], bx.lo[2]);
    const float3 hi = make_float3(bx.hi[0], bx.hi[1], bx.hi[2]);
    if (!hl_x_sphere_aabb(c.vpos, rc, lo, hi)) return false;
    if (c.cone.w != kNoCone && !hl_x_cone_aabb(c.vpos, c.cone, pad, lo, hi)) return false;
    return true;
}

// One wavefront per light.  The light bounds the cells it can reach -- the
// slices its view distance +- (radius + padding) spans and the columns whose
// lateral planes it can pass at its own depth, each widened by one -- and the
// wave tests those candidate cells 64 at a time with the exact membership
// rule, setting the member bits in the two-level bitmask.
__device__ __forceinline__ void mark_light(const ClusterArgs& a,
                                           const GpuCameraUniforms* __restrict__ cam,
                                           const CullLight* __restrict__ cl,
                                           const ColumnGeom* __restrict__ geom,
                                           const CellBox* __restrict__ boxes,
                                           uint32_t* __restrict__ bits,
                                           uint32_t* __restrict__ summary, uint32_t W, uint32_t SW,
                                           uint32_t li) {
    const uint32_t lane = threadIdx.x & 63u;
    if (li >= a.num_lights) return;
    const int gx = (int)a.grid_x, gy = (int)a.grid_y, gz = (int)a.grid_z;
    CullLight L = cl[li];
    L.vpos = make_float4(hl_uniform(L.vpos.x), hl_uniform(L.vpos.y), hl_uniform(L.vpos.z),
                         hl_uniform(L.vpos.w));
    L.cone = make_float4(hl_uniform(L.cone.x), hl_uniform(L.cone.y), hl_uniform(L.cone.z),
                         hl_uniform(L.cone.w));
    int x0 = 0, x1 = gx - 1, y0 = 0, y1 = gy - 1, s0 = 0, s1 = gz - 1;
    if (!__builtin_isinf(L.vpos.w)) {
        const float zc = -L.vpos.z;
        // slices under the largest padding, then again under the largest
        // padding of those slices (paddings grow with distance)
        float rr = L.vpos.w + a.r_pad_column;
        s1 = clampi(lower_bound(a.slice_near, gz, zc + rr) + 1, 0, gz - 1);
        rr = L.vpos.w + a.slice_r_pad[s1];
        s0 = lower_bound(a.slice_far, gz, zc - rr) - 1;
        s1 = lower_bound(a.slice_near, gz, zc + rr) + 1;
        if (zc > 1e-4f) {
            // A lateral plane through the origin at angle alpha from the view
            // axis passes a centre up to rr / cos(alpha) beyond the column's
            // edge at the centre's own depth; 1/cos(alpha) <= sqrt(1 + 1/p00^2).
            const float p00 = cam->proj[0], p11 = cam->proj[5];
            const float fx = sqrtf(1.0f + 1.0f / (p00 * p00));
            const float fy = sqrtf(1.0f + 1.0f / (p11 * p11));
            const float iz = 1.0f / zc;
            const float nx0 = p00 * (L.vpos.x - rr * fx) * iz, nx1 = p00 * (L.vpos.x + rr * fx) * iz;
            const float ny0 = p11 * (L.vpos.y - rr * fy) * iz, ny1 = p11 * (L.vpos.y + rr * fy) * iz;
            const float fw = (float)a.width, fh = (float)a.height;
            const float px0 = (nx0 + 1.0f) * 0.5f * fw, px1 = (nx1 + 1.0f) * 0.5f * fw;
            const float py0 = (1.0f - ny1) * 0.5f * fh, py1 = (1.0f - ny0) * 0.5f * fh;
            const float cw = (float)a.cell_w, ch = (float)a.cell_h;
            x0 = (int)fmaxf(floorf(px0 / cw) - 1.0f, -1.0f);
            x1 = (int)fminf(floorf(px1 / cw) + 1.0f, (float)gx);
            y0 = (int)fmaxf(floorf(py0 / ch) - 1.0f, -1.0f);
            y1 = (int)fminf(floorf(py1 / ch) + 1.0f, (float)gy);
        }
    }
    x0 = clampi(x0, 0, gx - 1);
    x1 = clampi(x1, -1, gx - 1);
    y0 = clampi(y0, 0, gy - 1);
    y1 = clampi(y1, -1, gy - 1);
    s0 = clampi(s0, 0, gz - 1);
    s1 = clampi(s1, -1, gz - 1);
    if (x1 < x0 || y1 < y0 || s1 < s0) return;
    const uint32_t nx = (uint32_t)(x1 - x0 + 1), ny = (uint32_t)(y1 - y0 + 1),
                   nz = (uint32_t)(s1 - s0 + 1);
    const uint32_t work = nx * ny * nz;
    const uint32_t ncol = a.grid_x * a.grid_y;
    const uint32_t bit = 1u << (li & 31u);
    const uint32_t word = li >> 5;
    for (uint32_t t0 = 0; t0 < work; t0 += 64u) {
        const uint32_t t = t0 + lane;
        if (t >= work) break;
        const uint32_t kz = t % nz;
        const uint32_t rest = t / nz;
        const uint32_t col = ((uint32_t)x0 + rest % nx) + a.grid_x * ((uint32_t)y0 + rest / nx);
        const uint32_t k = (uint32_t)s0 + kz;
        const ColumnGeom g = geom[col];
        if (!g.on) continue;
        const size_t cell = (size_t)col + (size_t)ncol * k;
        if (!cell_member(a, L, g, boxes[cell], k)) continue;
        const uint32_t old = atomicOr(&bits[cell * W + word], bit);
        if (old == 0u) atomicOr(&summary[cell * SW + (word >> 5)], 1u << (word & 31u));
    }
}

// one wavefront per cluster: number of members (before the cap)
__device__ __forceinline__ void count_cell(uint32_t ncell, uint32_t W, uint32_t SW,
                                           const uint32_t* __restrict__ bits,
                                           const uint32_t* __restrict__ summary,
                                           uint32_t* __restrict__ raw, uint32_t cell) {
    if (cell >= ncell) return;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t total = 0;
    for (uint32_t s0 = 0; s0 < SW; s0 += 64u) {
        const uint32_t j = s0 + lane;
        uint32_t sw = j < SW ? summary[(size_t)cell * SW + j] : 0u;
        while (sw) {
            const uint32_t b = (uint32_t)__builtin_ctz(sw);
            sw &= sw - 1u;
            total += (uint32_t)__popc(bits[(size_t)cell * W + 32u * j + b]);
        }
    }
    const uint32_t t = (uint32_t)dpp_sum((float)total);  // exact below 2^24 lights
    if (lane == 0) raw[cell] = t;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = hl_lane();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}

// Single block of 1024: capped counts -> offsets (exclusive scan in cluster
// order), pool truncation, overflow flags.  Tiles of 4096 counts, read
// coalesced, four per thread.
struct ScanSmem {
    unsigned long long wave[16];
    unsigned long long carry;
    uint32_t flags;
};

__device__ __forceinline__ void scan_cells(const uint32_t* __restrict__ raw, uint32_t ncell,
                                           uint32_t cap, uint32_t pool_cap,
                                           uint32_t* __restrict__ offsets,
                                           uint32_t* __restrict__ counts,
                                           uint32_t* __restrict__ flags_out, ScanSmem& sm) {
    unsigned long long* s_wave = sm.wave;
    unsigned long long& s_carry = sm.carry;
    uint32_t& s_flags = sm.flags;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    if (t == 0) {
        s_carry = 0;
        s_flags = 0;
    }
    __syncthreads();
    uint32_t fl = 0;
    for (uint32_t start = 0; start < ncell; start += 4096u) {
        uint32_t v[4];
        uint32_t local = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t c = start + 4u * t + (uint32_t)q;
            uint32_t n = c < ncell ? raw[c] : 0u;
            if (cap && n > cap) {
                n = cap;
                fl |= HL_OVERFLOW_CAP;
            }
            v[q] = n;
            local += n;
        }
        // local <= 4 * 2^24: the wave scan stays in 32 bits; tile sums go to 64
        const uint32_t incl = wave_incl_scan(local);
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        unsigned long long before = s_carry;
        for (uint32_t w = 0; w < wave; ++w) before += s_wave[w];
        unsigned long long off = before + (incl - local);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t c = start + 4u * t + (uint32_t)q;
            if (c < ncell) {
                const uint32_t o = off > pool_cap ? pool_cap : (uint32_t)off;
                const uint32_t room = pool_cap - o;
                uint32_t n = v[q];
                if (n > room) {
                    n = room;
                    fl |= HL_OVERFLOW_POOL;
                }
                offsets[c] = o;
                counts[c] = n;
            }
            off += v[q];
        }
        __syncthreads();
        if (t == 0) {
            unsigned long long tot = 0;
            for (uint32_t w = 0; w < 16u; ++w) tot += s_wave[w];
            s_carry += tot;
        }
        __syncthreads();
    }
    if (fl) atomicOr(&s_flags, fl);
    __syncthreads();
    if (t == 0) flags_out[0] |= s_flags;
}

// one wavefront per cluster: write the members in ascending order up to the
// (capped, pool-truncated) count, and clear the cluster's bitmask words.
__device__ __forceinline__ void fill_cell(uint32_t ncell, uint32_t W, uint32_t SW,
                                          uint32_t* __restrict__ bits,
                                          uint32_t* __restrict__ summary,
                                          const uint32_t* __restrict__ offsets,
                                          const uint32_t* __restrict__ counts,
                                          uint32_t* __restrict__ pool, uint32_t cell) {
    if (cell >= ncell) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t cnt = counts[cell];
    uint32_t* dst = pool + offsets[cell];
    uint32_t* cb = bits + (size_t)cell * W;
    uint32_t written = 0;
    for (uint32_t s0 = 0; s0 < SW; s0 += 64u) {
        const uint32_t j = s0 + lane;
        const uint32_t sw0 = j < SW ? summary[(size_t)cell * SW + j] : 0u;
        uint32_t lt = 0;
        for (uint32_t sw = sw0; sw; sw &= sw - 1u)
            lt += (uint32_t)__popc(cb[32u * j + (uint32_t)__builtin_ctz(sw)]);
        const uint32_t incl = wave_incl_scan(lt);
        uint32_t pos = written + incl - lt;
        for (uint32_t sw = sw0; sw; sw &= sw - 1u) {
            const uint32_t w = 32u * j + (uint32_t)__builtin_ctz(sw);
            uint32_t word = cb[w];
            cb[w] = 0u;
            while (word) {
                const uint32_t b = (uint32_t)__builtin_ctz(word);
                word &= word - 1u;
                if (pos < cnt) dst[pos] = 32u * w + b;
                ++pos;
            }
        }
        if (sw0) summary[(size_t)cell * SW + j] = 0u;
        written += hl_readlane(incl, 63);
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ------------------------------------------------------ one view per launch
__global__ __launch_bounds__(256) void k_cluster_geom(ClusterArgs a,
                                                      const GpuCameraUniforms* __restrict__ cam,
                                                      ColumnGeom* __restrict__ geom,
                                                      CellBox* __restrict__ boxes) {
    geom_cell(a, cam, geom, boxes, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_cluster_mark(ClusterArgs a,
                                                      const GpuCameraUniforms* __restrict__ cam,
                                                      const CullLight* __restrict__ cl,
                                                      const ColumnGeom* __restrict__ geom,
                                                      const CellBox* __restrict__ boxes,
                                                      uint32_t* __restrict__ bits,
                                                      uint32_t* __restrict__ summary, uint32_t W,
                                                      uint32_t SW) {
    mark_light(a, cam, cl, geom, boxes, bits, summary, W, SW, blockIdx.x * 4u + (threadIdx.x >> 6));
}

__global__ __launch_bounds__(256) void k_cluster_count(uint32_t ncell, uint32_t W, uint32_t SW,
                                                       const uint32_t* __restrict__ bits,
                                                       const uint32_t* __restrict__ summary,
                                                       uint32_t* __restrict__ raw) {
    count_cell(ncell, W, SW, bits, summary, raw, blockIdx.x * 4u + (threadIdx.x >> 6));
}

__global__ __launch_bounds__(1024) void k_cluster_scan(const uint32_t* __restrict__ raw,
                                                       uint32_t ncell, uint32_t cap,
                                                       uint32_t pool_cap,
                                                       uint32_t* __restrict__ offsets,
                                                       uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ flags_out) {
    __shared__ ScanSmem sm;
    scan_cells(raw, ncell, cap, pool_cap, offsets, counts, flags_out, sm);
}

__global__ __launch_bounds__(256) void k_cluster_fill(uint32_t ncell, uint32_t W, uint32_t SW,
                                                      uint32_t* __restrict__ bits,
                                                      uint32_t* __restrict__ summary,
                                                      const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ pool) {
    fill_cell(ncell, W, SW, bits, summary, offsets, counts, pool, blockIdx.x * 4u + (threadIdx.x >> 6));
}

// ------------------------------------------------- all views in one launch
// base[] holds nviews + 1 prefix offsets; returns the view owning index g.
__device__ __forceinline__ uint32_t find_view(const uint32_t* __restrict__ base, uint32_t nviews,
                                              uint32_t g) {
    uint32_t lo = 0, hi = nviews;  // base[lo] <= g < base[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ CellBox* view_boxes(const BatchCullView& v) {
    return (CellBox*)((uint8_t*)v.geom + cluster_geom_offset_boxes_dev(v.ca.grid_x * v.ca.grid_y));
}

__global__ __launch_bounds__(256) void k_batch_prep(const GpuLight* __restrict__ pool,
                                                    const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ light_base,
                                                    uint32_t nviews) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= light_base[nviews]) return;
    const uint32_t vi = find_view(light_base, nviews, g);
    const BatchCullView& v = views[vi];
    prep_one(pool, v.subset, v.ca.num_lights, v.cam, v.hot, v.cull, nullptr, g - light_base[vi]);
}

__global__ __launch_bounds__(256) void k_batch_geom(const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ cell_base,
                                                    uint32_t nviews) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= cell_base[nviews]) return;
    const uint32_t vi = find_view(cell_base, nviews, g);
    const BatchCullView& v = views[vi];
    geom_cell(v.ca, v.cam, (ColumnGeom*)v.geom, view_boxes(v), g - cell_base[vi]);
}

__global__ __launch_bounds__(256) void k_batch_mark(const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ light_base,
                                                    uint32_t nviews) {
    const uint32_t g = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (g >= light_base[nviews]) return;
    const uint32_t vi = find_view(light_base, nviews, g);
    const BatchCullView& v = views[vi];
    mark_light(v.ca, v.cam, v.cull, (const ColumnGeom*)v.geom, view_boxes(v), v.bits, v.summary,
               v.W, v.SW, g - light_base[vi]);
}

__global__ __launch_bounds__(256) void k_batch_count(const BatchCullView* __restrict__ views,
                                                     const uint32_t* __restrict__ cell_base,
                                                     uint32_t nviews) {
    const uint32_t g = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (g >= cell_base[nviews]) return;
    const uint32_t vi = find_view(cell_base, nviews, g);
    const BatchCullView& v = views[vi];
    count_cell(v.ncell, v.W, v.SW, v.bits, v.summary, v.raw, g - cell_base[vi]);
}

// one block per view
__global__ __launch_bounds__(1024) void k_batch_scan(const BatchCullView* __restrict__ views) {
    __shared__ ScanSmem sm;
    const BatchCullView& v = views[blockIdx.x];
    scan_cells(v.raw, v.ncell, v.ca.cap, v.pool_cap, v.offsets, v.counts, v.flags, sm);
}

__global__ __launch_bounds__(256) void k_batch_fill(const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ cell_base,
                                                    uint32_t nviews) {
    const uint32_t g = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (g >= cell_base[nviews]) return;
    const uint32_t vi = find_view(cell_base, nviews, g);
    const BatchCullView& v = views[vi];
    fill_cell(v.ncell, v.W, v.SW, v.bits, v.summary, v.offsets, v.counts, v.pool, g - cell_base[vi]);
}

hipError_t launch_batch_cull(hipStream_t s, const GpuLight* pool, const BatchCullView* views,
                             const uint32_t* light_base, const uint32_t* cell_base, uint32_t nviews,
                             uint32_t total_lights, uint32_t total_cells) {
    hipError_t e;
    if (total_lights) {
        k_batch_prep<<<(total_lights + 255) / 256, 256, 0, s>>>(pool, views, light_base, nviews);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    k_batch_geom<<<(total_cells + 255) / 256, 256, 0, s>>>(views, cell_base, nviews);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (total_lights) {
        k_batch_mark<<<(total_lights + 3) / 4, 256, 0, s>>>(views, light_base, nviews);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    k_batch_count<<<(total_cells + 3) / 4, 256, 0, s>>>(views, cell_base, nviews);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_batch_scan<<<nviews, 1024, 0, s>>>(views);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_batch_fill<<<(total_cells + 3) / 4, 256, 0, s>>>(views, cell_base, nviews);
    return hipGetLastError();
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

size_t cluster_geom_offset_boxes(uint32_t ncol) { return cluster_geom_offset_boxes_dev(ncol); }
size_t cluster_geom_bytes(uint32_t ncol, uint32_t ncell) {
    return cluster_geom_offset_boxes(ncol) + sizeof(CellBox) * (size_t)ncell;
}

hipError_t launch_cluster_cull(hipStream_t s, const ClusterArgs& a, const GpuCameraUniforms* cam,
                               const CullLight* cl, const ClusterScratch& w, uint32_t* offsets,
                               uint32_t* counts, uint32_t* pool, uint32_t pool_cap,
                               uint32_t* flags) {
    const uint32_t ncol = a.grid_x * a.grid_y;
    const uint32_t ncell = ncol * a.grid_z;
    const uint32_t W = (a.num_lights + 31u) / 32u, SW = (W + 31u) / 32u;
    ColumnGeom* geom = (ColumnGeom*)w.geom;
    CellBox* boxes = (CellBox*)((uint8_t*)w.geom + cluster_geom_offset_boxes(ncol));
    k_cluster_geom<<<(ncell + 255) / 256, 256, 0, s>>>(a, cam, geom, boxes);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.num_lights) {
        k_cluster_mark<<<(a.num_lights + 3) / 4, 256, 0, s>>>(a, cam, cl, geom, boxes, w.bits,
                                                                  w.summary, W, SW);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const uint32_t blocks = (ncell + 3) / 4;
    k_cluster_count<<<blocks, 256, 0, s>>>(ncell, W, SW, w.bits, w.summary, w.raw);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_cluster_scan<<<1, 1024, 0, s>>>(w.raw, ncell, a.cap, pool_cap, offsets, counts, flags);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_cluster_fill<<<blocks, 256, 0, s>>>(ncell, W, SW, w.bits, w.summary, offsets, counts, pool);
    return hipGetLastError();
}

}  // namespace hl
