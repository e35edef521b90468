// hl_cull.hip -- light preparation and light assignment kernels for gfx950.
//
//  k_prep_lights      GpuLight[] -> HotLight[] (shading) + CullLight[] (culling)
//  k_cull_tiles       reference 16-px tile lists (light_cull.wgsl:123-180), one
//                     wavefront per tile, ballot + prefix-popcount compaction
//  k_cluster_geom     per column: NDC rectangle and lateral planes; per cell:
//                     the view-space AABB; its blocks past the cells prepare
//                     the lights (as k_prep_lights) in the same launch
//  k_cluster_mark     light-parallel: each light bounds the cells it can reach
//                     (projected bounding box x slice range); a wavefront tests
//                     those candidates 64 at a time with the exact membership
//                     rule and sets the member bits in a two-level bitmask
//                     (bits + a summary bit per non-empty word); past its first
//                     chunks a light queues the rest
//  k_cluster_mark_tail the queued chunks, spread over all CUs
//  k_cluster_count    per cluster (16 lanes): popcount over its non-empty
//                     words; sums of the capped counts per scan tile of 1024
//  k_cluster_fill     per block of 16 clusters: its offsets from the tile
//                     sums and the counts before it (the exclusive scan, no
//                     separate launch), then the ascending indices from the
//                     bitmask, which it clears again for the next call
//  k_batch_*          the same steps for many views in one launch each
// Membership is decided by the same exact FP32 tests whatever the order of
// evaluation, so the lists equal the per-cluster brute force bit for bit.
#include "hl_internal.hpp"

namespace hl {

// ------------------------------------------------------------------- prep
// The culling record of a light (view-space centre and range, cone terms).
__device__ __forceinline__ CullLight cull_light(float4 pr, float4 dro, uint32_t type, float range,
                                                const float* __restrict__ v) {
    CullLight c;
    c.vpos = make_float4(hl_x_mat_row(v, 0, pr.x, pr.y, pr.z, 1.0f), hl_x_mat_row(v, 1, pr.x, pr.y, pr.z, 1.0f),
                         hl_x_mat_row(v, 2, pr.x, pr.y, pr.z, 1.0f), range);
    c.cone = make_float4(hl_x_mat_row(v, 0, dro.x, dro.y, dro.z, 0.0f), hl_x_mat_row(v, 1, dro.x, dro.y, dro.z, 0.0f),
                         hl_x_mat_row(v, 2, dro.x, dro.y, dro.z, 0.0f), type == HL_LIGHT_SPOT ? dro.w : kNoCone);
    c.axis = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    c.aux = c.axis;
    if (c.cone.w != kNoCone && !hl_x_cone_light(c.cone, &c.axis, &c.aux.x)) c.cone.w = kNoCone;
    return c;
}

__device__ __forceinline__ CullLight cull_light_at(const GpuLight* __restrict__ lights, uint32_t i,
                                                   const GpuCameraUniforms* __restrict__ cam) {
    const GpuLight* l = &lights[i];
    const float4 pr = *reinterpret_cast<const float4*>(l->position_range);
    const float4 dro = *reinterpret_cast<const float4*>(l->direction_outer);
    const uint32_t type = l->light_type;
    return cull_light(pr, dro, type, type == HL_LIGHT_DIRECTIONAL ? __builtin_inff() : pr.w, cam->view);
}

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
    if (hot)
        hot[i] = make_hot_light(pr.x, pr.y, pr.z, dro.x, dro.y, dro.z, ci, pr.w, dro.w, inner, type,
                                shadow_word(i, type, l->shadow_index), true);
    if (cull) cull[i] = cull_light(pr, dro, type, range, cam->view);
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

// The adjoint's set-up in one launch: the shading records, the per-light
// gradient accumulators cleared, and the loss (when given) cleared.
__global__ __launch_bounds__(256) void k_prep_zero(const GpuLight* __restrict__ lights, uint32_t n,
                                                   const GpuCameraUniforms* __restrict__ cam,
                                                   HotLight* __restrict__ hot, uint32_t* __restrict__ flags,
                                                   double* __restrict__ acc, size_t acc_n,
                                                   double* __restrict__ loss) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    prep_one(lights, nullptr, n, cam, hot, nullptr, flags, (uint32_t)(i < n ? i : n));
    for (size_t j = i; j < acc_n; j += (size_t)gridDim.x * blockDim.x) acc[j] = 0.0;
    if (i == 0 && loss) *loss = 0.0;
}

// -------------------------------------------------------------- tile cull
// The four planes of a 16x16 tile (light_cull.wgsl:123-150).
__device__ __forceinline__ void tile_planes(const GpuCameraUniforms* __restrict__ cam, const LightCullParams& p,
                                           uint32_t tile_x, uint32_t tile_y, float3 pl[4]) {
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

__device__ __forceinline__ bool tile_pass(const float3 pl[4], float4 c) {
    return hl_x_plane_pass(pl[0], c, c.w) && hl_x_plane_pass(pl[1], c, c.w) && hl_x_plane_pass(pl[2], c, c.w) &&
           hl_x_plane_pass(pl[3], c, c.w);
}

// One wavefront per 16x16 tile; four tiles per 256-thread block; every light
// tested (any tile grid).
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
    tile_planes(cam, p, tile_x, tile_y, pl);
    uint32_t count = 0;
    uint32_t* out = lists + (size_t)tile * HL_MAX_LIGHTS_PER_TILE;
    for (uint32_t base = 0; base < p.num_lights; base += 64u) {
        const uint32_t i = base + lane;
        bool pass = false;
        if (i < p.num_lights) pass = tile_pass(pl, cl[i].vpos);
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
// chunk_fetch reads both records as words at these offsets
static_assert(sizeof(float3) == 12 && sizeof(ColumnGeom) == 80 && offsetof(ColumnGeom, on) == 64,
              "ColumnGeom layout");
static_assert(sizeof(CellBox) == 24, "CellBox layout");

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
// The records of cell (col, k): its view-space box, and with planes its
// column's record.
__device__ __forceinline__ void cell_geometry(const ClusterArgs& a, const GpuCameraUniforms* __restrict__ cam,
                                              uint32_t col, uint32_t k, bool planes, ColumnGeom& g,
                                              CellBox& bx) {
#pragma clang fp contract(off)
    const uint32_t cx = col % a.grid_x, cy = col / a.grid_x;
    float l = 0.0f, r = 0.0f, b = 0.0f, t = 0.0f;
    const bool on = column_rect(a, cx, cy, &l, &r, &b, &t);
    const float p00 = cam->proj[0], p11 = cam->proj[5];
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
    if (!planes) return;
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
}

__device__ __forceinline__ void geom_cell(const ClusterArgs& a,
                                          const GpuCameraUniforms* __restrict__ cam,
                                          ColumnGeom* __restrict__ geom,
                                          CellBox* __restrict__ boxes, uint32_t cell) {
    const uint32_t ncol = a.grid_x * a.grid_y;
    if (cell >= ncol * a.grid_z) return;
    const uint32_t col = cell % ncol, k = cell / ncol;
    ColumnGeom g;
    CellBox bx;
    cell_geometry(a, cam, col, k, k == 0, g, bx);
    boxes[cell] = bx;
    if (k == 0) geom[col] = g;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Membership of light (c) in cell (col, k): the definition of the clustered
// lists -- lateral planes of the column with the column padding, the cell's
// AABB with the slice padding (pad = slice_r_pad[k]), and the cone for spot
// lights.  Every test is evaluated (no short circuit) so that the column and
// box records are loaded together rather than one after another.
__device__ __forceinline__ bool cell_member(const CullLight& c, const ColumnGeom& g,
                                            const CellBox& bx, float pad) {
#pragma clang fp contract(off)
    const float rc = c.vpos.w + pad;
    const float3 lo = make_float3(bx.lo[0], bx.lo[1], bx.lo[2]);
    const float3 hi = make_float3(bx.hi[0], bx.hi[1], bx.hi[2]);
    bool in = g.on != 0u;
    in &= hl_x_plane_pass(g.pl[0], c.vpos, rc);
    in &= hl_x_plane_pass(g.pl[1], c.vpos, rc);
    in &= hl_x_plane_pass(g.pl[2], c.vpos, rc);
    in &= hl_x_plane_pass(g.pl[3], c.vpos, rc);
    in &= hl_x_sphere_aabb(c.vpos, rc, lo, hi);
    if (c.cone.w != kNoCone) in &= hl_x_cone_cell(c.vpos, c.axis, c.aux.x, pad, lo, hi);
    return in;
}

// Candidate cells of a light: the columns and slices left by necessary
// conditions of the membership rule, rounded outwards.
//  * slices: the box test needs slice_near[k] <= zc + rc and slice_far[k] >=
//    zc - rc (zc the light's view distance, rc = |radius| + pad_k, at most rr,
//    the radius plus the largest padding of the slices in reach: paddings do
//    not decrease with k);
//  * columns: a column with NDC edges l < r (|l|, |r| <= 1) passes its left
//    plane (normal (1, 0, l / p00) normalised) only if
//    l <= p00 (cx + rc sqrt(1 + l^2 / p00^2)) / zc <= p00 (cx + rr fx) / zc
//    with fx = sqrt(1 + 1 / p00^2), and its right plane only if
//    r >= p00 (cx - rr fx) / zc (zc > 0, p00 > 0); in pixels, the columns
//    floor(px0 / cell_w) .. floor(px1 / cell_w).  Rows alike with p11.
// The FP32 tests can pass a few ulp beyond these real-number bounds, and the
// bounds are computed with fast reciprocals and square roots (a few ulp of the
// terms): rr is widened by 1e-5 of itself and of |cx| + |cy| + zc, the pixel
// bounds by 1e-5 of their magnitude plus a quarter pixel, the slice limits by
// 1e-5 relative -- each many times those errors.  Directional lights, lights
// with a non-finite centre and a camera with p00 or p11 <= 0 take the whole
// grid (or the whole screen); the exact tests decide there.
struct MarkBox {
    uint32_t x0, y0, s0, nx, ny, nz;
};

// All 64 lanes must be active (the slice limits are found by ballots); sp is
// slice_r_pad[lane] (lanes below grid_z; slice_near, slice_far and the
// paddings are non-decreasing, checked by the host).
__device__ __forceinline__ bool light_cells(const ClusterArgs& a,
                                            const GpuCameraUniforms* __restrict__ cam,
                                            const CullLight& L, float sp, MarkBox& m) {
    const int gx = (int)a.grid_x, gy = (int)a.grid_y, gz = (int)a.grid_z;
    const uint32_t lane = hl_lane();
    const bool sl = (int)lane < gz;
    const float vn = sl ? a.slice_near[lane] : 0.0f, vf = sl ? a.slice_far[lane] : 0.0f;
    const float sn = vn - 1e-5f * fabsf(vn), sf = vf + 1e-5f * fabsf(vf);
    int x0 = 0, x1 = gx - 1, y0 = 0, y1 = gy - 1, s0 = 0, s1 = gz - 1;
    const float zc = -L.vpos.z;
    if (!__builtin_isinf(L.vpos.w) && __builtin_isfinite(zc) && __builtin_isfinite(L.vpos.x) &&
        __builtin_isfinite(L.vpos.y)) {
        const float slack = 1e-5f * (fabsf(L.vpos.x) + fabsf(L.vpos.y) + fabsf(zc));
        const float r = fabsf(L.vpos.w);
        // slices in reach under the largest padding, then under the largest
        // padding of those slices
        float rr = (r + hl_readlane(sp, (uint32_t)(gz - 1))) * (1.0f + 1e-5f) + slack;
        s1 = (int)__popcll(hl_ballot(sl && sn <= zc + rr)) - 1;
        if (s1 < 0) return false;
        rr = (r + hl_readlane(sp, (uint32_t)s1)) * (1.0f + 1e-5f) + slack;
        s0 = (int)__popcll(hl_ballot(sl && sf < zc - rr));
        s1 = (int)__popcll(hl_ballot(sl && sn <= zc + rr)) - 1;
        const float p00 = cam->proj[0], p11 = cam->proj[5];
        if (zc > 1e-4f && p00 > 0.0f && p11 > 0.0f) {
            const float fx = __builtin_amdgcn_sqrtf(1.0f + hl_f_rcp(p00 * p00));
            const float fy = __builtin_amdgcn_sqrtf(1.0f + hl_f_rcp(p11 * p11));
            const float iz = hl_f_rcp(zc);
            const float nx0 = p00 * (L.vpos.x - rr * fx) * iz, nx1 = p00 * (L.vpos.x + rr * fx) * iz;
            const float ny0 = p11 * (L.vpos.y - rr * fy) * iz, ny1 = p11 * (L.vpos.y + rr * fy) * iz;
            const float fw = (float)a.width, fh = (float)a.height;
            float px0 = (nx0 + 1.0f) * 0.5f * fw, px1 = (nx1 + 1.0f) * 0.5f * fw;
            float py0 = (1.0f - ny1) * 0.5f * fh, py1 = (1.0f - ny0) * 0.5f * fh;
            px0 -= 1e-5f * fabsf(px0) + 0.25f;
            px1 += 1e-5f * fabsf(px1) + 0.25f;
            py0 -= 1e-5f * fabsf(py0) + 0.25f;
            py1 += 1e-5f * fabsf(py1) + 0.25f;
            const float icw = hl_f_rcp((float)a.cell_w), ich = hl_f_rcp((float)a.cell_h);
            // clamped to [-1, g] before the conversion (NaN: the whole range)
            x0 = (int)fminf(fmaxf(floorf(px0 * icw), -1.0f), (float)gx);
            x1 = (int)fmaxf(fminf(floorf(px1 * icw), (float)gx), -1.0f);
            y0 = (int)fminf(fmaxf(floorf(py0 * ich), -1.0f), (float)gy);
            y1 = (int)fmaxf(fminf(floorf(py1 * ich), (float)gy), -1.0f);
        }
    }
    x0 = clampi(x0, 0, gx - 1);
    x1 = clampi(x1, -1, gx - 1);
    y0 = clampi(y0, 0, gy - 1);
    y1 = clampi(y1, -1, gy - 1);
    s0 = clampi(s0, 0, gz - 1);
    s1 = clampi(s1, -1, gz - 1);
    if (x1 < x0 || y1 < y0 || s1 < s0) return false;
    m.x0 = (uint32_t)x0;
    m.y0 = (uint32_t)y0;
    m.s0 = (uint32_t)s0;
    m.nx = (uint32_t)(x1 - x0 + 1);
    m.ny = (uint32_t)(y1 - y0 + 1);
    m.nz = (uint32_t)(s1 - s0 + 1);
    return true;
}

__device__ __forceinline__ uint4 pack_box(const MarkBox& m) {
    return make_uint4(m.x0, m.y0, (m.s0 << 16) | m.nz, (m.nx << 16) | m.ny);
}
__device__ __forceinline__ MarkBox unpack_box(uint4 v) {
    MarkBox m;
    m.x0 = v.x;
    m.y0 = v.y;
    m.s0 = v.z >> 16;
    m.nz = v.z & 0xffffu;
    m.nx = v.w >> 16;
    m.ny = v.w & 0xffffu;
    return m;
}

// Candidate cells chunk*64 .. chunk*64+63 of light li (one per lane, slice
// fastest): the exact membership rule, then the member bit and its summary
// bit (no-return atomics).  Split in a fetch (the lane's cell and its column,
// box and slice-padding records, loaded together) and a test-and-mark, so
// that a wave can have the loads of two chunks in flight at once.
struct ChunkCell {
    ColumnGeom g;
    CellBox bx;
    size_t cell;
    float pad;
    bool on;
};

__device__ __forceinline__ ChunkCell chunk_fetch(const ClusterArgs& a, const MarkBox& m,
                                                 const ColumnGeom* __restrict__ geom,
                                                 const CellBox* __restrict__ boxes, uint32_t chunk) {
    ChunkCell r;
    const uint32_t work = m.nx * m.ny * m.nz;
    const uint32_t t = chunk * 64u + (threadIdx.x & 63u);
    r.on = t < work;
    if (!r.on) return r;
    uint32_t rest, iy;
    if (work <= (1u << 21)) {
        rest = hl_floor_div(t, m.nz);
        iy = hl_floor_div(rest, m.nx);
    } else {
        rest = t / m.nz;
        iy = rest / m.nx;
    }
    const uint32_t kz = t - rest * m.nz, ix = rest - iy * m.nx;
    const uint32_t col = (m.x0 + ix) + a.grid_x * (m.y0 + iy);
    const uint32_t k = m.s0 + kz;
    r.cell = (size_t)col + (size_t)(a.grid_x * a.grid_y) * k;
    // records of a batched view come through pointers read from memory:
    // hl_g keeps their loads global rather than flat
    {
        const hl_gptr<const hl_f32x4> gq = hl_g(reinterpret_cast<const hl_f32x4*>(geom + col));
        const hl_f32x4 q0 = gq[0], q1 = gq[1], q2 = gq[2], q4 = gq[4];
        r.g.pl[0] = make_float3(q0.x, q0.y, q0.z);
        r.g.pl[1] = make_float3(q0.w, q1.x, q1.y);
        r.g.pl[2] = make_float3(q1.z, q1.w, q2.x);
        r.g.pl[3] = make_float3(q2.y, q2.z, q2.w);
        r.g.on = __float_as_uint(q4.x);
        const hl_gptr<const float> bq = hl_g(&boxes[r.cell].lo[0]);
        r.bx.lo[0] = bq[0];
        r.bx.lo[1] = bq[1];
        r.bx.lo[2] = bq[2];
        r.bx.hi[0] = bq[3];
        r.bx.hi[1] = bq[4];
        r.bx.hi[2] = bq[5];
    }
    r.pad = a.slice_r_pad[k];
    return r;
}

__device__ __forceinline__ size_t cluster_cells(const ClusterArgs& a) {
    return (size_t)a.grid_x * a.grid_y * a.grid_z;
}

// The bitmask is stored word-major: word w of cluster c (lights 32 w .. 32 w
// + 31) at bits[w * ncell + c], summary word s (bitmask words 32 s .. 32 s +
// 31) at summary[s * ncell + c].  A light's member cells in one row of one
// slice are then adjacent words (fewer lines for its atomics), and the list
// kernel's 16 consecutive clusters read their words of one light range from
// one line.
__device__ __forceinline__ void chunk_mark(const CullLight& L, const ChunkCell& r, uint32_t li,
                                           uint32_t* __restrict__ bits, uint32_t* __restrict__ summary,
                                           size_t ncell) {
    if (!r.on || !cell_member(L, r.g, r.bx, r.pad)) return;
    const uint32_t word = li >> 5;
    __hip_atomic_fetch_or(hl_g(&bits[(size_t)word * ncell + r.cell]), 1u << (li & 31u), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(hl_g(&summary[(size_t)(word >> 5) * ncell + r.cell]), 1u << (word & 31u),
                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void mark_chunk(const ClusterArgs& a, const CullLight& L,
                                           const MarkBox& m, uint32_t li,
                                           const ColumnGeom* __restrict__ geom,
                                           const CellBox* __restrict__ boxes,
                                           uint32_t* __restrict__ bits,
                                           uint32_t* __restrict__ summary, uint32_t chunk) {
    chunk_mark(L, chunk_fetch(a, m, geom, boxes, chunk), li, bits, summary, cluster_cells(a));
}

__device__ __forceinline__ CullLight load_light_uniform(const CullLight* __restrict__ cl, uint32_t li) {
    CullLight L = cl[li];
    L.vpos = make_float4(hl_uniform(L.vpos.x), hl_uniform(L.vpos.y), hl_uniform(L.vpos.z),
                         hl_uniform(L.vpos.w));
    L.cone = make_float4(hl_uniform(L.cone.x), hl_uniform(L.cone.y), hl_uniform(L.cone.z),
                         hl_uniform(L.cone.w));
    L.axis = make_float4(hl_uniform(L.axis.x), hl_uniform(L.axis.y), hl_uniform(L.axis.z),
                         hl_uniform(L.axis.w));
    L.aux.x = hl_uniform(L.aux.x);
    return L;
}

// One wavefront per light: its candidate box, its first kMarkInline chunks,
// and the rest queued as (qlight, qview << 24 | chunk) entries with the box
// stored at bounds[qlight] (qlight: the light's index in the queue's
// numbering, qview: its view in a batch; chunks stay below 2^20).
// Chunks that do not fit in the light's sub-queue stay with this wave.
__device__ __forceinline__ void mark_light(const ClusterArgs& a,
                                           const GpuCameraUniforms* __restrict__ cam,
                                           const CullLight* __restrict__ cl,
                                           const ColumnGeom* __restrict__ geom,
                                           const CellBox* __restrict__ boxes,
                                           uint32_t* __restrict__ bits,
                                           uint32_t* __restrict__ summary,
                                           uint32_t li, const MarkQueue& q, uint32_t qlight,
                                           uint32_t qview) {
    if (li >= a.num_lights) return;
    const CullLight L = load_light_uniform(cl, li);
    const float sp = hl_lane() < a.grid_z ? a.slice_r_pad[hl_lane()] : 0.0f;
    MarkBox m;
    if (!light_cells(a, cam, L, sp, m)) return;
    const uint32_t nchunk = (m.nx * m.ny * m.nz + 63u) / 64u;
    // chunks 0 and 1 always stay with this wave: their loads go out first,
    // both together, ahead of the queueing of the rest
    static_assert(kMarkInline == 2, "two head chunks fetched together");
    const ChunkCell r0 = chunk_fetch(a, m, geom, boxes, 0u);
    ChunkCell r1;
    if (nchunk > 1u) r1 = chunk_fetch(a, m, geom, boxes, 1u);
    uint32_t qlo = nchunk, qhi = nchunk;  // chunks [qlo, qhi) are queued
    if (nchunk > kMarkInline) {
        const uint32_t extra = nchunk - kMarkInline;
        const uint32_t sub = qlight % kMarkQueues;
        uint32_t base = 0;
        if ((threadIdx.x & 63u) == 0) base = atomicAdd(q.count + sub * kMarkCountStride, extra);
        base = hl_readlane(base, 0);
        // room left in the sub-queue, saturating at zero.  Written as
        // fit = base >= seg ? 0 : min(extra, seg - base), -O3 marked the
        // difference no-wrap and folded the select away, and an overflowing
        // light then took fit = extra and wrote past its segment
        // (tests/test_mark_queue_guard.py checks the compiled form)
        const uint32_t room = __builtin_elementwise_sub_sat(q.seg, base);
        const uint32_t fit = min(extra, room);
        if (fit) {
            if ((threadIdx.x & 63u) == 0) q.bounds[qlight] = pack_box(m);
            uint2* seg = q.entries + (size_t)sub * q.seg + base;
            for (uint32_t j = threadIdx.x & 63u; j < fit; j += 64u)
                seg[j] = make_uint2(qlight, (qview << 24) | (kMarkInline + j));
            qlo = kMarkInline;
            qhi = kMarkInline + fit;
        }
    }
    // chunks [0, qlo) and [qhi, nchunk) stay here (qlo >= 2 whenever nchunk >= 2)
    chunk_mark(L, r0, li, bits, summary, cluster_cells(a));
    if (nchunk > 1u) chunk_mark(L, r1, li, bits, summary, cluster_cells(a));
    for (uint32_t c = 2u; c < qlo; ++c) mark_chunk(a, L, m, li, geom, boxes, bits, summary, c);
    for (uint32_t c = qhi; c < nchunk; ++c) mark_chunk(a, L, m, li, geom, boxes, bits, summary, c);
}

// Count and fill work on 16 lanes per cluster (four clusters per wavefront).
// The non-empty bitmask words of a cluster are first listed in LDS in
// ascending order (lane j of a group owns summary word j (+ 16 r), i.e. bitmask
// words 32 j .. 32 j + 31) and then read together, eight per lane, so that a
// cluster costs one round trip to memory for its words whatever their number
// (up to kWordWin per window).
constexpr uint32_t kCellLanes = 16;
constexpr uint32_t kWordWin = 128;
constexpr uint32_t kWordsPerLane = kWordWin / kCellLanes;
struct CellSmem {
    uint32_t words[256 / kCellLanes][kWordWin];
};

__device__ __forceinline__ uint32_t group_incl_scan(uint32_t v) {
    const uint32_t sl = threadIdx.x & (kCellLanes - 1u);
#pragma unroll
    for (uint32_t o = 1; o < kCellLanes; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, kCellLanes);
        if (sl >= o) v += u;
    }
    return v;
}

// Calls f(list, n) for each window of up to kWordWin non-empty word indices of
// cluster cell (when on), in ascending order, uniformly over the cluster's 16
// lanes (SW is uniform over them).  clear: zero the summary words afterwards.
template <class F>
__device__ __forceinline__ void for_nonempty_words(uint32_t* __restrict__ summary, size_t ncell,
                                                   uint32_t cell, uint32_t SW, bool on, bool clear,
                                                   uint32_t* list, F&& f) {
    const uint32_t sl = threadIdx.x & (kCellLanes - 1u);
    for (uint32_t s0 = 0; s0 < SW; s0 += kCellLanes) {
        const uint32_t j = s0 + sl;
        const bool in = on && j < SW;
        uint32_t* const sp = summary + (in ? (size_t)j * ncell + cell : 0);
        const uint32_t sw = in ? *sp : 0u;
        const uint32_t c = (uint32_t)__popc(sw);
        const uint32_t incl = group_incl_scan(c);
        const uint32_t T = __shfl(incl, kCellLanes - 1u, kCellLanes);
        const uint32_t P = incl - c;
        for (uint32_t wb = 0; wb < T; wb += kWordWin) {
            uint32_t k = P;
            for (uint32_t m = sw; m; m &= m - 1u, ++k)
                if (k >= wb && k < wb + kWordWin) list[k - wb] = 32u * j + (uint32_t)__builtin_ctz(m);
            hl_wave_sync();
            f(list, T - wb < kWordWin ? T - wb : kWordWin);
            hl_wave_sync();
        }
        if (clear && sw) *sp = 0u;
    }
}

// number of members of a cluster (before the cap), returned to every lane of
// the cluster's group; lane 0 of the group also stores it in raw[cell]
__device__ __forceinline__ uint32_t count_cell(uint32_t ncell, uint32_t SW,
                                               const uint32_t* __restrict__ bits,
                                               const uint32_t* __restrict__ summary,
                                               uint32_t* __restrict__ raw, uint32_t cell,
                                               CellSmem& sm) {
    const uint32_t sl = threadIdx.x & (kCellLanes - 1u);
    const bool on = cell < ncell;
    const uint32_t* cb = bits + (on ? cell : 0u);  // word w at cb[w * ncell]
    uint32_t total = 0;
    for_nonempty_words(const_cast<uint32_t*>(summary), ncell, cell, SW, on, false,
                       sm.words[threadIdx.x / kCellLanes], [&](const uint32_t* list, uint32_t n) {
                           uint32_t wv[kWordsPerLane];
#pragma unroll
                           for (uint32_t q = 0; q < kWordsPerLane; ++q) {
                               const uint32_t e = kWordsPerLane * sl + q;
                               wv[q] = cb[(size_t)list[e < n ? e : 0u] * ncell];
                           }
#pragma unroll
                           for (uint32_t q = 0; q < kWordsPerLane; ++q)
                               if (kWordsPerLane * sl + q < n) total += (uint32_t)__popc(wv[q]);
                       });
#pragma unroll
    for (uint32_t o = kCellLanes / 2; o >= 1; o >>= 1) total += __shfl_xor(total, o, kCellLanes);
    if (raw && on && sl == 0) raw[cell] = total;
    return total;
}

// Per block of 16 clusters: adds each cluster's capped count to the sum of
// its scan tile (tile_sum null: no cluster), one atomic per run of equal
// tiles (a block's clusters lie in one tile except at view boundaries of a
// batch).
struct TileSumSmem {
    unsigned long long v[256 / kCellLanes];
    unsigned long long* p[256 / kCellLanes];
};
__device__ __forceinline__ void add_tile_sums(TileSumSmem& sm, unsigned long long* tile_sum,
                                              uint32_t capped) {
    if ((threadIdx.x & (kCellLanes - 1u)) == 0) {
        sm.v[threadIdx.x / kCellLanes] = capped;
        sm.p[threadIdx.x / kCellLanes] = tile_sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long* cur = nullptr;
        unsigned long long acc = 0;
        for (uint32_t i = 0; i < 256u / kCellLanes; ++i) {
            if (sm.p[i] != cur) {
                if (cur && acc) atomicAdd(cur, acc);
                cur = sm.p[i];
                acc = 0;
            }
            acc += sm.v[i];
        }
        if (cur && acc) atomicAdd(cur, acc);
    }
}

constexpr uint32_t kScanVPT = kScanTile / 256u;  // clusters per thread when a block reads a scan tile

// members of a cluster in ascending order, the first cnt of them written to
// dst (its capped, pool-truncated count); clears the cluster's bitmask words
// for the next call
__device__ __forceinline__ void fill_cell_at(uint32_t ncell, uint32_t SW, uint32_t* __restrict__ bits,
                                             uint32_t* __restrict__ summary, uint32_t cell,
                                             uint32_t cnt, uint32_t* __restrict__ dst, CellSmem& sm) {
    const uint32_t sl = threadIdx.x & (kCellLanes - 1u);
    const bool on = cell < ncell;
    uint32_t* cb = bits + (on ? cell : 0u);  // word w at cb[w * ncell]
    uint32_t written = 0;
    for_nonempty_words(summary, ncell, cell, SW, on, true,
                       sm.words[threadIdx.x / kCellLanes], [&](const uint32_t* list, uint32_t n) {
                           uint32_t wi[kWordsPerLane], wv[kWordsPerLane];
#pragma unroll
                           for (uint32_t q = 0; q < kWordsPerLane; ++q) {
                               const uint32_t e = kWordsPerLane * sl + q;
                               wi[q] = list[e < n ? e : 0u];
                               wv[q] = cb[(size_t)wi[q] * ncell];
                           }
                           uint32_t lt = 0;
#pragma unroll
                           for (uint32_t q = 0; q < kWordsPerLane; ++q) {
                               if (kWordsPerLane * sl + q >= n) wv[q] = 0u;
                               lt += (uint32_t)__popc(wv[q]);
                           }
                           const uint32_t incl = group_incl_scan(lt);
                           uint32_t pos = written + incl - lt;
#pragma unroll
                           for (uint32_t q = 0; q < kWordsPerLane; ++q) {
                               if (kWordsPerLane * sl + q >= n) continue;
                               cb[(size_t)wi[q] * ncell] = 0u;
                               for (uint32_t m = wv[q]; m; m &= m - 1u, ++pos)
                                   if (pos < cnt) dst[pos] = 32u * wi[q] + (uint32_t)__builtin_ctz(m);
                           }
                           written += __shfl(incl, kCellLanes - 1u, kCellLanes);
                       });
}

__device__ __forceinline__ void fill_cell(uint32_t ncell, uint32_t SW,
                                          uint32_t* __restrict__ bits,
                                          uint32_t* __restrict__ summary,
                                          const uint32_t* __restrict__ offsets,
                                          const uint32_t* __restrict__ counts,
                                          uint32_t* __restrict__ pool, uint32_t cell, CellSmem& sm) {
    const bool on = cell < ncell;
    fill_cell_at(ncell, SW, bits, summary, cell, on ? counts[cell] : 0u, on ? pool + offsets[cell] : pool, sm);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ------------------------------------------------------ scan fused into fill
// A fill block (16 consecutive clusters, in one scan tile of kScanTile) finds
// its clusters' offsets itself: the sums of the preceding scan tiles (from the
// count kernel), the capped raw counts of the tile's clusters before the
// block, and the prefix within the block -- at most 27 + 1008 values at 4K,
// read in one round trip -- then the pool truncation and overflow flags, then
// the lists.  The separate scan launch is gone.
struct ScanFillSmem {
    uint32_t capped[256 / kCellLanes];
    unsigned long long part[4];
    uint32_t flags;
};

__device__ __forceinline__ void scan_fill_block(const uint32_t* __restrict__ raw, uint32_t ncell,
                                                uint32_t SW, uint32_t* __restrict__ bits,
                                                uint32_t* __restrict__ summary, uint32_t cap,
                                                uint32_t pool_cap, uint32_t* __restrict__ offsets,
                                                uint32_t* __restrict__ counts, uint32_t* __restrict__ pool,
                                                uint32_t* __restrict__ flags_out,
                                                uint32_t* __restrict__ flags_out2, uint32_t first,
                                                CellSmem& cs, ScanFillSmem& sm) {
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t grp = t / kCellLanes, sl = t & (kCellLanes - 1u);
    if (first >= ncell) return;  // a batched view's padding cells (uniform)
    const unsigned long long* tile_sums =
        (const unsigned long long*)((const uint8_t*)raw + cluster_tile_sums_offset(ncell));
    const uint32_t tile = first / kScanTile, t0 = tile * kScanTile;
    if (t == 0) sm.flags = 0u;
    // sums of the preceding tiles, and the tile's clusters [t0, first)
    unsigned long long acc = 0;
    for (uint32_t j = t; j < tile; j += 256u) acc += tile_sums[j];
    const uint32_t c0 = t0 + kScanVPT * t;  // 4 clusters per thread, 16-B aligned
    if (c0 < first) {
        const uint4 u = *reinterpret_cast<const uint4*>(raw + c0);  // first <= ncell, a multiple of 16
        const uint32_t v[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) acc += (cap && v[q] > cap) ? cap : v[q];
    }
    acc = wave_sum_u64(acc);
    if (lane == 0) sm.part[wave] = acc;
    // the block's own clusters
    const uint32_t cell = first + grp;
    const bool on = cell < ncell;
    uint32_t n = on ? raw[cell] : 0u;
    uint32_t fl = 0;
    if (cap && n > cap) {
        n = cap;
        fl |= HL_OVERFLOW_CAP;
    }
    if (sl == 0) sm.capped[grp] = n;
    __syncthreads();
    unsigned long long off = sm.part[0] + sm.part[1] + sm.part[2] + sm.part[3];
    for (uint32_t j = 0; j < grp; ++j) off += sm.capped[j];
    const uint32_t o = off > pool_cap ? pool_cap : (uint32_t)off;
    uint32_t cnt = n;
    if (cnt > pool_cap - o) {
        cnt = pool_cap - o;
        fl |= HL_OVERFLOW_POOL;
    }
    if (on && sl == 0) {
        offsets[cell] = o;
        counts[cell] = cnt;
    }
    fill_cell_at(ncell, SW, bits, summary, cell, cnt, pool + o, cs);
    if (fl && sl == 0) atomicOr(&sm.flags, fl);
    __syncthreads();
    if (t == 0 && sm.flags) {
        atomicOr(flags_out, sm.flags);
        if (flags_out2) atomicOr(flags_out2, sm.flags);
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ------------------------------------------------------ one view per launch
// The geometry kernel also resets the mark queue, the overflow words and the
// scan tile sums (it runs before every kernel that reads them).  Its blocks past the cells'
// prepare the lights' records (independent work that the mark kernel needs
// too), so that both take one launch.
__global__ __launch_bounds__(256) void k_cluster_geom(ClusterArgs a,
                                                      const GpuCameraUniforms* __restrict__ cam,
                                                      ColumnGeom* __restrict__ geom,
                                                      CellBox* __restrict__ boxes,
                                                      uint32_t* __restrict__ qcount,
                                                      uint32_t* __restrict__ flags,
                                                      uint32_t* __restrict__ flags2,
                                                      unsigned long long* __restrict__ tile_sums,
                                                      uint32_t ntiles,
                                                      const GpuLight* __restrict__ lights,
                                                      HotLight* __restrict__ hot,
                                                      CullLight* __restrict__ cl,
                                                      uint32_t* __restrict__ feature_flags,
                                                      uint32_t geom_blocks) {
    if (blockIdx.x >= geom_blocks) {
        prep_one(lights, nullptr, a.num_lights, cam, hot, cl, feature_flags,
                 (blockIdx.x - geom_blocks) * blockDim.x + threadIdx.x);
        return;
    }
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < kMarkQueues) qcount[g * kMarkCountStride] = 0u;
    if (g == 0) {
        *flags = 0u;
        if (flags2) *flags2 = 0u;
    }
    if (g < ntiles) tile_sums[g] = 0ull;
    geom_cell(a, cam, geom, boxes, g);
}

__global__ __launch_bounds__(256) void k_cluster_mark(ClusterArgs a,
                                                      const GpuCameraUniforms* __restrict__ cam,
                                                      const CullLight* __restrict__ cl,
                                                      const ColumnGeom* __restrict__ geom,
                                                      const CellBox* __restrict__ boxes,
                                                      uint32_t* __restrict__ bits,
                                                      uint32_t* __restrict__ summary, MarkQueue q) {
    const uint32_t li = blockIdx.x * 4u + (threadIdx.x >> 6);
    mark_light(a, cam, cl, geom, boxes, bits, summary, li, q, li, 0u);
}

// Queued chunks over a fixed grid of wavefronts: nw / kMarkQueues waves per
// sub-queue stride over its valid entries, min(count, seg), then exit.
__global__ __launch_bounds__(256) void k_cluster_mark_tail(ClusterArgs a,
                                                           const CullLight* __restrict__ cl,
                                                           const ColumnGeom* __restrict__ geom,
                                                           const CellBox* __restrict__ boxes,
                                                           uint32_t* __restrict__ bits,
                                                           uint32_t* __restrict__ summary, MarkQueue q) {
    const uint32_t w = blockIdx.x * 4u + (threadIdx.x >> 6), nw = gridDim.x * 4u;
    const uint32_t sub = w % kMarkQueues, step = nw / kMarkQueues;  // nw: a multiple of kMarkQueues
    const uint32_t n = min(hl_uniform(q.count[sub * kMarkCountStride]), q.seg);
    const uint2* seg = q.entries + (size_t)sub * q.seg;
    for (uint32_t e = w / kMarkQueues; e < n; e += step) {
        const uint2 en = seg[e];
        const uint32_t li = hl_uniform(en.x), chunk = hl_uniform(en.y) & 0xFFFFFFu;
        const CullLight L = load_light_uniform(cl, li);
        const uint4 pb = q.bounds[li];
        const MarkBox m = unpack_box(make_uint4(hl_uniform(pb.x), hl_uniform(pb.y), hl_uniform(pb.z),
                                                hl_uniform(pb.w)));
        mark_chunk(a, L, m, li, geom, boxes, bits, summary, chunk);
    }
}

__global__ __launch_bounds__(256) void k_cluster_count(uint32_t ncell, uint32_t SW,
                                                       const uint32_t* __restrict__ bits,
                                                       const uint32_t* __restrict__ summary,
                                                       uint32_t* __restrict__ raw, uint32_t cap) {
    __shared__ TileSumSmem sm;
    __shared__ CellSmem cs;
    const uint32_t cell = (blockIdx.x * blockDim.x + threadIdx.x) / kCellLanes;
    const uint32_t n = count_cell(ncell, SW, bits, summary, raw, cell, cs);
    unsigned long long* ts = (unsigned long long*)((uint8_t*)raw + cluster_tile_sums_offset(ncell));
    add_tile_sums(sm, cell < ncell ? ts + cell / kScanTile : nullptr, (cap && n > cap) ? cap : n);
}

__global__ __launch_bounds__(256) void k_cluster_fill(uint32_t ncell, uint32_t SW, const uint32_t* __restrict__ raw,
                                                      uint32_t* __restrict__ bits,
                                                      uint32_t* __restrict__ summary, uint32_t cap,
                                                      uint32_t pool_cap, uint32_t* __restrict__ offsets,
                                                      uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ pool,
                                                      uint32_t* __restrict__ flags_out,
                                                      uint32_t* __restrict__ flags_out2) {
    __shared__ CellSmem cs;
    __shared__ ScanFillSmem sm;
    scan_fill_block(raw, ncell, SW, bits, summary, cap, pool_cap, offsets, counts, pool, flags_out, flags_out2,
                    blockIdx.x * (256u / kCellLanes), cs, sm);
}

// ------------------------------------------------- all views in one launch
// base[] holds nviews + 1 prefix offsets; returns the view owning index g.
// Cell ranges start on multiples of kBatchCellAlign, so a block of the
// geometry, count and fill kernels lies in one view: its lookup (and every
// view field it reads) depends on blockIdx only and stays scalar.
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

// ------------------------------------------------- clusters, brute force
// Small problems (cells x lights <= kBruteTests) take two launches instead of
// the light-parallel chain: every cell tests every light with the membership
// rule (cell_member), one wave per cell and four cells per block, the
// lights' culling records built per block into LDS kBfLights at a time, each
// cell's records built by its own wave in registers.  k_cluster_bf_count
// stores each cell's member count (and the lights' shading records);
// k_cluster_bf_fill finds each cell's offset from the capped counts before
// it, then writes the ascending list.  Membership does not depend on the
// order of evaluation, so the lists are the light-parallel path's.
constexpr uint64_t kBruteTests = 1ull << 22;
constexpr uint32_t kBfCells = 4;  // one per wave
constexpr uint32_t kBfLights = 256;

struct BfSmem {
    CullLight light[kBfLights];
    uint32_t capped[kBfCells];
    unsigned long long part[4];
    uint32_t flags;
};

// The records of the wave's cell (every lane alike).
struct BfCell {
    ColumnGeom g;
    CellBox bx;
    float pad;
};
__device__ __forceinline__ BfCell bf_cell(const ClusterArgs& a, const GpuCameraUniforms* __restrict__ cam,
                                          uint32_t cell, uint32_t ncell) {
    BfCell c;
    const uint32_t ncol = a.grid_x * a.grid_y;
    if (cell < ncell) {
        cell_geometry(a, cam, cell % ncol, cell / ncol, true, c.g, c.bx);
        c.pad = a.slice_r_pad[cell / ncol];
    } else {  // no cell: no member (the records are not read)
        c.g = ColumnGeom{};
        c.bx = CellBox{};
        c.pad = 0.0f;
    }
    return c;
}

// For each light of the LDS chunk [l0, l0 + kBfLights) in ascending order, in
// groups of 64: f(group, pass) with pass this lane's light's membership in
// the wave's cell (uniform calls).
template <class F>
__device__ __forceinline__ void bf_chunk(const ClusterArgs& a, uint32_t l0, const BfCell& c, const BfSmem& sm,
                                         F&& f) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nl = a.num_lights - l0 < kBfLights ? a.num_lights - l0 : kBfLights;
    for (uint32_t j0 = 0; j0 < nl; j0 += 64u) {
        const uint32_t j = j0 + lane;
        const bool pass = c.g.on != 0u && j < nl && cell_member(sm.light[j], c.g, c.bx, c.pad);
        f(j0, pass);
    }
}

__device__ __forceinline__ void bf_load_lights(const ClusterArgs& a, const GpuCameraUniforms* __restrict__ cam,
                                               const GpuLight* __restrict__ lights, uint32_t l0, BfSmem& sm) {
    __syncthreads();  // the previous chunk is no longer read
    const uint32_t i = l0 + threadIdx.x;
    if (i < a.num_lights) sm.light[threadIdx.x] = cull_light_at(lights, i, cam);
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_cluster_bf_count(ClusterArgs a, const GpuCameraUniforms* __restrict__ cam,
                                                          const GpuLight* __restrict__ lights,
                                                          HotLight* __restrict__ hot,
                                                          uint32_t* __restrict__ feature_flags,
                                                          uint32_t* __restrict__ raw, uint32_t* __restrict__ flags,
                                                          uint32_t* __restrict__ flags2) {
    __shared__ BfSmem sm;
    const uint32_t ncell = a.grid_x * a.grid_y * a.grid_z;
    const uint32_t cell = blockIdx.x * kBfCells + (threadIdx.x >> 6);
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // OR-ed by the fill launch
        *flags = 0u;
        if (flags2) *flags2 = 0u;
    }
    // the shading records, grid-stride over the lights
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.num_lights; i += gridDim.x * blockDim.x)
        prep_one(lights, nullptr, a.num_lights, cam, hot, nullptr, feature_flags, i);
    const BfCell c = bf_cell(a, cam, cell, ncell);
    uint32_t cnt = 0u;
    for (uint32_t l0 = 0; l0 < a.num_lights; l0 += kBfLights) {
        bf_load_lights(a, cam, lights, l0, sm);
        bf_chunk(a, l0, c, sm, [&](uint32_t, bool pass) { cnt += (uint32_t)__popcll(hl_ballot(pass)); });
    }
    if ((threadIdx.x & 63u) == 0 && cell < ncell) raw[cell] = cnt;
}

__global__ __launch_bounds__(256) void k_cluster_bf_fill(ClusterArgs a, const GpuCameraUniforms* __restrict__ cam,
                                                         const GpuLight* __restrict__ lights,
                                                         const uint32_t* __restrict__ raw, uint32_t pool_cap,
                                                         uint32_t* __restrict__ offsets, uint32_t* __restrict__ counts,
                                                         uint32_t* __restrict__ pool, uint32_t* __restrict__ flags,
                                                         uint32_t* __restrict__ flags2) {
    __shared__ BfSmem sm;
    const uint32_t ncell = a.grid_x * a.grid_y * a.grid_z, first = blockIdx.x * kBfCells;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    const uint32_t cell = first + wave, cap = a.cap;
    // capped counts of the cells before the block
    unsigned long long acc = 0;
    for (uint32_t j = t; j < first; j += blockDim.x) {
        const uint32_t v = raw[j];
        acc += (cap && v > cap) ? cap : v;
    }
    acc = wave_sum_u64(acc);
    if (lane == 0) sm.part[wave] = acc;
    if (t == 0) sm.flags = 0u;
    if (t < kBfCells) {
        const uint32_t v = first + t < ncell ? raw[first + t] : 0u;
        sm.capped[t] = (cap && v > cap) ? cap : v;
    }
    const BfCell c = bf_cell(a, cam, cell, ncell);
    __syncthreads();
    // the wave's cell: offset, count after the cap and the pool truncation
    unsigned long long base = sm.part[0] + sm.part[1] + sm.part[2] + sm.part[3];
    for (uint32_t j = 0; j < wave; ++j) base += sm.capped[j];
    const uint32_t o = base > pool_cap ? pool_cap : (uint32_t)base;
    uint32_t cnt = sm.capped[wave], fl = 0u;
    if (cell < ncell && cap && raw[cell] > cap) fl |= HL_OVERFLOW_CAP;
    if (cnt > pool_cap - o) {
        cnt = pool_cap - o;
        fl |= HL_OVERFLOW_POOL;
    }
    if (lane == 0 && cell < ncell) {
        offsets[cell] = o;
        counts[cell] = cnt;
        if (fl) atomicOr(&sm.flags, fl);
    }
    uint32_t written = 0u;
    for (uint32_t l0 = 0; l0 < a.num_lights; l0 += kBfLights) {
        bf_load_lights(a, cam, lights, l0, sm);
        bf_chunk(a, l0, c, sm, [&](uint32_t j0, bool pass) {
            const uint64_t m = hl_ballot(pass);
            const uint32_t pos = written + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (pass && pos < cnt) pool[o + pos] = l0 + j0 + lane;
            written += (uint32_t)__popcll(m);
        });
    }
    __syncthreads();
    if (t == 0 && sm.flags) {
        atomicOr(flags, sm.flags);
        if (flags2) atomicOr(flags2, sm.flags);
    }
}

// Cells of every view, then (blocks past geom_blocks) the lights of every

// view gathered from the pool, in one launch.
__global__ __launch_bounds__(256) void k_batch_geom(const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ cell_base,
                                                    uint32_t nviews, uint32_t* __restrict__ qcount,
                                                    const GpuLight* __restrict__ pool,
                                                    const uint32_t* __restrict__ light_base,
                                                    uint32_t geom_blocks) {
    // the views' overflow words start each frame at zero (nviews <= 255)
    if (blockIdx.x == 0 && threadIdx.x < nviews) *views[threadIdx.x].flags = 0u;
    if (blockIdx.x >= geom_blocks) {
        const uint32_t l = (blockIdx.x - geom_blocks) * blockDim.x + threadIdx.x;
        if (l >= light_base[nviews]) return;
        const uint32_t vi = find_view(light_base, nviews, l);
        const BatchCullView& v = views[vi];
        prep_one(pool, v.subset, v.ca.num_lights, v.cam, v.hot, v.cull, nullptr, l - light_base[vi]);
        return;
    }
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < kMarkQueues) qcount[g * kMarkCountStride] = 0u;
    const uint32_t vi = find_view(cell_base, nviews, blockIdx.x * blockDim.x);
    const BatchCullView& v = views[vi];
    const uint32_t c = g - cell_base[vi];
    if (c < cluster_scan_tiles_dev(v.ncell))
        ((unsigned long long*)((uint8_t*)v.raw + cluster_tile_sums_offset(v.ncell)))[c] = 0ull;
    geom_cell(v.ca, v.cam, (ColumnGeom*)v.geom, view_boxes(v), c);
}

// one wavefront per light of any view; queue entries carry the global index
__global__ __launch_bounds__(256) void k_batch_mark(const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ light_base,
                                                    uint32_t nviews, MarkQueue q) {
    const uint32_t g = hl_uniform(blockIdx.x * 4u + (threadIdx.x >> 6));
    if (g >= light_base[nviews]) return;
    const uint32_t vi = find_view(light_base, nviews, g);
    const BatchCullView& v = views[vi];
    mark_light(v.ca, v.cam, v.cull, (const ColumnGeom*)v.geom, view_boxes(v), v.bits, v.summary,
               g - light_base[vi], q, g, vi);
}

__global__ __launch_bounds__(256) void k_batch_mark_tail(const BatchCullView* __restrict__ views,
                                                         const uint32_t* __restrict__ light_base,
                                                         uint32_t nviews, MarkQueue q) {
    const uint32_t w = blockIdx.x * 4u + (threadIdx.x >> 6), nw = gridDim.x * 4u;
    const uint32_t sub = w % kMarkQueues, step = nw / kMarkQueues;  // nw: a multiple of kMarkQueues
    const uint32_t n = min(hl_uniform(q.count[sub * kMarkCountStride]), q.seg);
    const uint2* seg = q.entries + (size_t)sub * q.seg;
    for (uint32_t e = w / kMarkQueues; e < n; e += step) {
        const uint2 en = seg[e];
        const uint32_t g = hl_uniform(en.x), packed = hl_uniform(en.y);
        const uint32_t vi = packed >> 24, chunk = packed & 0xFFFFFFu;
        const BatchCullView& v = views[vi];
        const uint32_t li = g - light_base[vi];
        const CullLight L = load_light_uniform(v.cull, li);
        const uint4 pb = q.bounds[g];
        const MarkBox m = unpack_box(make_uint4(hl_uniform(pb.x), hl_uniform(pb.y), hl_uniform(pb.z),
                                                hl_uniform(pb.w)));
        mark_chunk(v.ca, L, m, li, (const ColumnGeom*)v.geom, view_boxes(v), v.bits, v.summary, chunk);
    }
}

__global__ __launch_bounds__(256) void k_batch_count(const BatchCullView* __restrict__ views,
                                                     const uint32_t* __restrict__ cell_base,
                                                     uint32_t nviews) {
    __shared__ TileSumSmem sm;
    __shared__ CellSmem cs;
    constexpr uint32_t kCellsPerBlock = 256u / kCellLanes;
    const uint32_t vi = find_view(cell_base, nviews, blockIdx.x * kCellsPerBlock);
    const BatchCullView& v = views[vi];
    const uint32_t c = blockIdx.x * kCellsPerBlock + threadIdx.x / kCellLanes - cell_base[vi];
    const uint32_t n = count_cell(v.ncell, v.SW, v.bits, v.summary, v.raw, c, cs);
    unsigned long long* ts = (unsigned long long*)((uint8_t*)v.raw + cluster_tile_sums_offset(v.ncell));
    const uint32_t cap = v.ca.cap;
    add_tile_sums(sm, c < v.ncell ? ts + c / kScanTile : nullptr, (cap && n > cap) ? cap : n);
}

// scan and fill of 16 clusters of one view (scan_fill_block)
__global__ __launch_bounds__(256) void k_batch_fill(const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ cell_base,
                                                    uint32_t nviews) {
    constexpr uint32_t kCellsPerBlock = 256u / kCellLanes;
    const uint32_t vi = find_view(cell_base, nviews, blockIdx.x * kCellsPerBlock);
    const BatchCullView& v = views[vi];
    __shared__ CellSmem cs;
    __shared__ ScanFillSmem sm;
    scan_fill_block(v.raw, v.ncell, v.SW, v.bits, v.summary, v.ca.cap, v.pool_cap, v.offsets, v.counts, v.pool,
                    v.flags, nullptr, blockIdx.x * kCellsPerBlock - cell_base[vi], cs, sm);
}

// grid of the queued-chunk launches: 64 wavefronts per sub-queue, enough to
// spread the chunks of a few heavy lights over every CU whatever the light
// count (idle waves read one counter and exit); a multiple of kMarkQueues / 4
static uint32_t tail_blocks(uint32_t) { return 16u * kMarkQueues; }

hipError_t launch_batch_cull(hipStream_t s, const GpuLight* pool, const BatchCullView* views,
                             const uint32_t* light_base, const uint32_t* cell_base,
                             const uint32_t* tile_base, uint32_t nviews, uint32_t total_lights,
                             uint32_t total_cells, uint32_t total_tiles, const MarkQueue& mq) {
    if (total_cells % kBatchCellAlign) return hipErrorInvalidValue;
    hipError_t e;
    const uint32_t geom_blocks = total_cells / 256u;  // total_cells: kBatchCellAlign multiple
    k_batch_geom<<<geom_blocks + (total_lights + 255) / 256, 256, 0, s>>>(views, cell_base, nviews, mq.count,
                                                                          pool, light_base, geom_blocks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (total_lights) {
        k_batch_mark<<<(total_lights + 3) / 4, 256, 0, s>>>(views, light_base, nviews, mq);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        k_batch_mark_tail<<<tail_blocks(total_lights), 256, 0, s>>>(views, light_base, nviews, mq);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const uint32_t cblocks = total_cells / (256u / kCellLanes);  // total_cells: kBatchCellAlign multiple
    k_batch_count<<<cblocks, 256, 0, s>>>(views, cell_base, nviews);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_batch_fill<<<cblocks, 256, 0, s>>>(views, cell_base, nviews);
    return hipGetLastError();
}

// ------------------------------------------------------ tile cull, binned
// Candidate tiles of a light for the tile planes' test (inclusive rectangle;
// false when empty).  Necessary conditions of passing a tile's four planes
// (normals (1, 0, l/p00), -(1, 0, r/p00), (0, -1, -t/p11), (0, 1, b/p11),
// normalised by s = sqrt(1 + e^2/p00^2) <= fx for an edge |e| <= 1), with
// the light's centre c, view distance zc = -c.z and radius r (|r| <= rr):
//  - in front (zc > 0): l <= p00 (cx + rr fx) / zc and r >= p00 (cx - rr fx) / zc,
//    rows alike with p11 (the cluster cull's bound);
//  - behind the camera (zc < 0, |zc| = -zc): the planes' half-spaces meet
//    only near the mirrored projection, l >= p00 (-cx - rr fx) / |zc| and
//    r <= p00 (rr fx - cx) / |zc|, t <= p11 (rr fy - cy) / |zc| and
//    b >= p11 (-rr fy - cy) / |zc|;
//  - within 1e-4 of the camera plane, a non-finite centre or radius, or a
//    camera with p00 or p11 <= 0: every tile.
// Rounded outwards as the cluster cull's candidates: rr widened by 1e-5 of
// itself and of |cx| + |cy| + |zc|, the pixel bounds by 1e-5 of their
// magnitude plus a quarter pixel.
__device__ __forceinline__ bool light_tiles(const GpuCameraUniforms* __restrict__ cam, const LightCullParams& p,
                                            float4 c, int* tx0, int* tx1, int* ty0, int* ty1) {
    const int ntx = (int)p.num_tiles_x, nty = (int)p.num_tiles_y;
    *tx0 = 0;
    *tx1 = ntx - 1;
    *ty0 = 0;
    *ty1 = nty - 1;
    const float zc = -c.z, p00 = cam->proj[0], p11 = cam->proj[5];
    if (!__builtin_isfinite(c.w) || !__builtin_isfinite(c.x) || !__builtin_isfinite(c.y) ||
        !__builtin_isfinite(zc) || !(p00 > 0.0f && p11 > 0.0f) || fabsf(zc) <= 1e-4f)
        return ntx > 0 && nty > 0;
    const float slack = 1e-5f * (fabsf(c.x) + fabsf(c.y) + fabsf(zc));
    const float rr = fabsf(c.w) * (1.0f + 1e-5f) + slack;
    const float fx = __builtin_amdgcn_sqrtf(1.0f + hl_f_rcp(p00 * p00));
    const float fy = __builtin_amdgcn_sqrtf(1.0f + hl_f_rcp(p11 * p11));
    const float iz = hl_f_rcp(fabsf(zc));
    float nx0, nx1, nt, nb;  // column edges: l <= nx1, r >= nx0 (front) or l >= nx0, r <= nx1 (behind)
    if (zc > 0.0f) {
        nx0 = p00 * (c.x - rr * fx) * iz;
        nx1 = p00 * (c.x + rr * fx) * iz;
        nb = p11 * (c.y - rr * fy) * iz;  // t >= nb
        nt = p11 * (c.y + rr * fy) * iz;  // b <= nt
    } else {
        nx0 = p00 * (-c.x - rr * fx) * iz;
        nx1 = p00 * (rr * fx - c.x) * iz;
        nt = p11 * (rr * fy - c.y) * iz;   // t <= nt
        nb = p11 * (-rr * fy - c.y) * iz;  // b >= nb
    }
    const float fw = (float)p.screen_width, fh = (float)p.screen_height;
    // in both cases the candidate pixel columns lie in [px0, px1] and rows in
    // [py0, py1]: front, from l <= nx1 and r >= nx0; behind, from l >= nx0 and
    // r <= nx1 (the same interval, reached from the other side)
    float px0 = (nx0 + 1.0f) * 0.5f * fw, px1 = (nx1 + 1.0f) * 0.5f * fw;
    float py0 = (1.0f - nt) * 0.5f * fh, py1 = (1.0f - nb) * 0.5f * fh;
    px0 -= 1e-5f * fabsf(px0) + 0.25f;
    px1 += 1e-5f * fabsf(px1) + 0.25f;
    py0 -= 1e-5f * fabsf(py0) + 0.25f;
    py1 += 1e-5f * fabsf(py1) + 0.25f;
    const float its = 1.0f / (float)HL_TILE_SIZE;  // exact
    // clamped to [-1, n] before the conversion (NaN: the whole range)
    const int x0 = (int)fminf(fmaxf(floorf(px0 * its), -1.0f), (float)ntx);
    const int x1 = (int)fmaxf(fminf(floorf(px1 * its), (float)ntx), -1.0f);
    const int y0 = (int)fminf(fmaxf(floorf(py0 * its), -1.0f), (float)nty);
    const int y1 = (int)fmaxf(fminf(floorf(py1 * its), (float)nty), -1.0f);
    *tx0 = clampi(x0, 0, ntx - 1);
    *tx1 = clampi(x1, -1, ntx - 1);
    *ty0 = clampi(y0, 0, nty - 1);
    *ty1 = clampi(y1, -1, nty - 1);
    return *tx1 >= *tx0 && *ty1 >= *ty0;
}

// One wavefront per light: the super-tiles its candidate tiles touch get
// its bit (word-major bitmask over nsup cells, as the cluster cull's).
// Block 0 also zeroes the scan tile sums that the count kernel adds to; the
// blocks past the lights' (plane_blocks of them) build each tile's four
// planes once, one thread per tile, for the list kernel's waves (a tile's
// planes cost its wave some 800 VALU instructions of exact divisions and
// square roots, the same in every lane).
__global__ __launch_bounds__(256) void k_tile_bin(const GpuCameraUniforms* __restrict__ cam, LightCullParams p,
                                                  const CullLight* __restrict__ cl, uint32_t* __restrict__ bits,
                                                  uint32_t* __restrict__ summary,
                                                  unsigned long long* __restrict__ tile_sums, uint32_t ntile_sums,
                                                  float* __restrict__ planes, uint32_t plane_blocks) {
    if (blockIdx.x < plane_blocks) {
        const uint32_t tile = blockIdx.x * blockDim.x + threadIdx.x;
        if (tile < p.num_tiles_x * p.num_tiles_y) {
            float3 pl[4];
            tile_planes(cam, p, tile % p.num_tiles_x, tile / p.num_tiles_x, pl);
            float* o = planes + 12 * (size_t)tile;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                o[3 * k] = pl[k].x;
                o[3 * k + 1] = pl[k].y;
                o[3 * k + 2] = pl[k].z;
            }
        }
        return;
    }
    if (blockIdx.x == plane_blocks)
        for (uint32_t i = threadIdx.x; i < ntile_sums; i += blockDim.x) tile_sums[i] = 0ull;
    const uint32_t li = hl_uniform((blockIdx.x - plane_blocks) * 4u + (threadIdx.x >> 6));
    if (li >= p.num_lights) return;
    const float4 c = cl[li].vpos;
    const float4 cu = make_float4(hl_uniform(c.x), hl_uniform(c.y), hl_uniform(c.z), hl_uniform(c.w));
    int tx0, tx1, ty0, ty1;
    if (!light_tiles(cam, p, cu, &tx0, &tx1, &ty0, &ty1)) return;
    const uint32_t sx = (p.num_tiles_x + kSuperTile - 1u) / kSuperTile;
    const uint32_t nsup = sx * ((p.num_tiles_y + kSuperTile - 1u) / kSuperTile);
    const uint32_t s0x = (uint32_t)tx0 / kSuperTile, s1x = (uint32_t)tx1 / kSuperTile;
    const uint32_t s0y = (uint32_t)ty0 / kSuperTile, s1y = (uint32_t)ty1 / kSuperTile;
    const uint32_t nx = s1x - s0x + 1u, work = nx * (s1y - s0y + 1u);
    const uint32_t word = li >> 5;
    for (uint32_t t = threadIdx.x & 63u; t < work; t += 64u) {
        const uint32_t iy = t / nx, ix = t - iy * nx;
        const size_t sup = (size_t)(s0x + ix) + (size_t)sx * (s0y + iy);
        __hip_atomic_fetch_or(hl_g(&bits[(size_t)word * nsup + sup]), 1u << (li & 31u), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_or(hl_g(&summary[(size_t)(word >> 5) * nsup + sup]), 1u << (word & 31u),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One wavefront per tile: its super-tile's candidates (ascending) through
// the exact plane test, the first HL_MAX_LIGHTS_PER_TILE kept in order.  The
// candidates go through 256 at a time, four per lane (lane l takes the l-th
// of each run of 64), their loads issued together.
__global__ __launch_bounds__(256) void k_tile_lists(const float* __restrict__ planes, LightCullParams p,
                                                    const CullLight* __restrict__ cl,
                                                    const uint32_t* __restrict__ soff,
                                                    const uint32_t* __restrict__ scnt,
                                                    const uint32_t* __restrict__ spool,
                                                    uint32_t* __restrict__ lists, uint32_t* __restrict__ counts) {
    constexpr uint32_t kPer = 4;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t total = p.num_tiles_x * p.num_tiles_y;
    if (tile >= total) return;
    const uint32_t tile_x = tile % p.num_tiles_x, tile_y = tile / p.num_tiles_x;
    const uint32_t sup = tile_x / kSuperTile + ((p.num_tiles_x + kSuperTile - 1u) / kSuperTile) * (tile_y / kSuperTile);
    float3 pl[4];
    {
        const float* q = planes + 12 * (size_t)tile;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            pl[k] = make_float3(hl_uniform(q[3 * k]), hl_uniform(q[3 * k + 1]), hl_uniform(q[3 * k + 2]));
    }
    const uint32_t n = hl_uniform(scnt[sup]);
    const uint32_t* cand = spool + hl_uniform(soff[sup]);
    uint32_t count = 0;
    uint32_t* out = lists + (size_t)tile * HL_MAX_LIGHTS_PER_TILE;
    for (uint32_t base = 0; base < n && count < HL_MAX_LIGHTS_PER_TILE; base += 64u * kPer) {
        uint32_t idx[kPer];
        float4 c[kPer];
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            const uint32_t j = base + 64u * q + lane;
            idx[q] = j < n ? cand[j] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q)
            c[q] = idx[q] != 0xFFFFFFFFu ? cl[idx[q]].vpos : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (uint32_t q = 0; q < kPer; ++q) {
            const bool pass = idx[q] != 0xFFFFFFFFu && tile_pass(pl, c[q]);
            const uint64_t m = hl_ballot(pass);
            const uint32_t pos = count + hl_prefix(m);
            if (pass && pos < HL_MAX_LIGHTS_PER_TILE) out[pos] = idx[q];
            count += (uint32_t)__popcll(m);
        }
    }
    if (lane == 0) counts[tile] = count < HL_MAX_LIGHTS_PER_TILE ? count : HL_MAX_LIGHTS_PER_TILE;
}

// ---------------------------------------------------------------- launchers
hipError_t launch_prep(hipStream_t s, const GpuLight* lights, const uint32_t* subset, uint32_t n,
                       const GpuCameraUniforms* cam, HotLight* hot, CullLight* cull,
                       uint32_t* flags) {
    if (n == 0) return hipSuccess;
    k_prep_lights<<<(n + 255) / 256, 256, 0, s>>>(lights, subset, n, cam, hot, cull, flags);
    return hipGetLastError();
}

hipError_t launch_prep_zero(hipStream_t s, const GpuLight* lights, uint32_t n, const GpuCameraUniforms* cam,
                            HotLight* hot, uint32_t* flags, double* acc, size_t acc_n, double* loss) {
    // one thread per light (prep_one), the accumulators grid-stride
    const size_t prep_blocks = ((size_t)n + 255) / 256;
    size_t zero_blocks = (acc_n + 255) / 256;
    if (zero_blocks > 4096) zero_blocks = 4096;
    size_t blocks = prep_blocks > zero_blocks ? prep_blocks : zero_blocks;
    if (blocks == 0) blocks = 1;
    k_prep_zero<<<(unsigned)blocks, 256, 0, s>>>(lights, n, cam, hot, flags, acc, acc_n, loss);
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

hipError_t launch_cull_tiles_binned(hipStream_t s, const GpuCameraUniforms* cam, const LightCullParams& p,
                                    const CullLight* cl, uint32_t* lists, uint32_t* counts,
                                    const TileBinScratch& w) {
    const uint32_t total = p.num_tiles_x * p.num_tiles_y;
    if (total == 0) return hipSuccess;
    const uint32_t nsup = tile_bin_supers(p);
    const uint32_t W = (p.num_lights + 31u) / 32u, SW = (W + 31u) / 32u;
    hipError_t e;
    const uint32_t plane_blocks = (total + 255u) / 256u;
    k_tile_bin<<<plane_blocks + (p.num_lights + 3) / 4, 256, 0, s>>>(
        cam, p, cl, w.bits, w.summary, (unsigned long long*)((uint8_t*)w.raw + cluster_tile_sums_offset(nsup)),
        cluster_scan_tiles(nsup), w.planes, plane_blocks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t blocks = (nsup + 256u / kCellLanes - 1u) / (256u / kCellLanes);
    k_cluster_count<<<blocks, 256, 0, s>>>(nsup, SW, w.bits, w.summary, w.raw, 0u);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // the pool holds every (super-tile, light) pair: nothing is truncated
    k_cluster_fill<<<blocks, 256, 0, s>>>(nsup, SW, w.raw, w.bits, w.summary, 0u, w.pool_cap, w.offsets,
                                          w.counts, w.pool, w.flags, nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_tile_lists<<<(total + 3) / 4, 256, 0, s>>>(w.planes, p, cl, w.offsets, w.counts, w.pool, lists, counts);
    return hipGetLastError();
}

size_t cluster_geom_offset_boxes(uint32_t ncol) { return cluster_geom_offset_boxes_dev(ncol); }
size_t cluster_geom_bytes(uint32_t ncol, uint32_t ncell) {
    return cluster_geom_offset_boxes(ncol) + sizeof(CellBox) * (size_t)ncell;
}

hipError_t launch_cluster_cull(hipStream_t s, const ClusterArgs& a, const GpuCameraUniforms* cam,
                               const GpuLight* lights, HotLight* hot, CullLight* cl, uint32_t* feature_flags,
                               const ClusterScratch& w, uint32_t* offsets,
                               uint32_t* counts, uint32_t* pool, uint32_t pool_cap,
                               uint32_t* flags, uint32_t* flags2) {
    const uint32_t ncol = a.grid_x * a.grid_y;
    const uint32_t ncell = ncol * a.grid_z;
    if ((uint64_t)ncell * a.num_lights <= kBruteTests) {
        const uint32_t blocks = (ncell + kBfCells - 1u) / kBfCells;
        k_cluster_bf_count<<<blocks, 256, 0, s>>>(a, cam, lights, hot, feature_flags, w.raw, flags, flags2);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        k_cluster_bf_fill<<<blocks, 256, 0, s>>>(a, cam, lights, w.raw, pool_cap, offsets, counts, pool, flags,
                                                 flags2);
        return hipGetLastError();
    }
    const uint32_t W = (a.num_lights + 31u) / 32u, SW = (W + 31u) / 32u;
    ColumnGeom* geom = (ColumnGeom*)w.geom;
    CellBox* boxes = (CellBox*)((uint8_t*)w.geom + cluster_geom_offset_boxes(ncol));
    const uint32_t ntiles = cluster_scan_tiles(ncell);
    unsigned long long* tile_sums = (unsigned long long*)((uint8_t*)w.raw + cluster_tile_sums_offset(ncell));
    const uint32_t geom_blocks = (ncell + 255) / 256;
    k_cluster_geom<<<geom_blocks + (a.num_lights + 255) / 256, 256, 0, s>>>(
        a, cam, geom, boxes, w.mq.count, flags, flags2, tile_sums, ntiles, lights, hot, cl, feature_flags,
        geom_blocks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.num_lights) {
        k_cluster_mark<<<(a.num_lights + 3) / 4, 256, 0, s>>>(a, cam, cl, geom, boxes, w.bits,
                                                                  w.summary, w.mq);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        k_cluster_mark_tail<<<tail_blocks(a.num_lights), 256, 0, s>>>(a, cl, geom, boxes, w.bits,
                                                                       w.summary, w.mq);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const uint32_t blocks = (ncell + 256u / kCellLanes - 1u) / (256u / kCellLanes);
    k_cluster_count<<<blocks, 256, 0, s>>>(ncell, SW, w.bits, w.summary, w.raw, a.cap);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_cluster_fill<<<blocks, 256, 0, s>>>(ncell, SW, w.raw, w.bits, w.summary, a.cap, pool_cap, offsets, counts,
                                          pool, flags, flags2);
    return hipGetLastError();
}

}  // namespace hl
