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
//  k_cluster_count    per cluster (16 lanes): popcount over its non-empty words
//  k_cluster_scan     exclusive scan of the capped counts, a block per tile of
//                     1024 clusters (tile sums come from the count kernel)
//  k_cluster_fill     per cluster (16 lanes): ascending indices from the
//                     bitmask, which it clears again for the next call
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
    if (hot)
        hot[i] = make_hot_light(pr.x, pr.y, pr.z, dro.x, dro.y, dro.z, ci, pr.w, dro.w, inner, type,
                                shadow_word(i, type, l->shadow_index), true);
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
        c.axis = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        c.aux = c.axis;
        if (c.cone.w != kNoCone && !hl_x_cone_light(c.cone, &c.axis, &c.aux.x)) c.cone.w = kNoCone;
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

// Candidate cells of a light: the slices its view distance +- (radius +
// padding) spans and the columns whose lateral planes it can pass at its own
// depth, each widened by one.  False when there are none.
struct MarkBox {
    uint32_t x0, y0, s0, nx, ny, nz;
};

// first k in [0, n) with v[k] >= x, for ascending v held one per lane (lane
// k holds v[k], n <= 64): the number of lanes below n with v[k] < x
__device__ __forceinline__ int lane_lower_bound(float vk, int n, float x) {
    return (int)__popcll(hl_ballot((int)hl_lane() < n && vk < x));
}

// All 64 lanes must be active (the slice bounds are found by ballots); sp is
// slice_r_pad[lane] (lanes below grid_z).
__device__ __forceinline__ bool light_cells(const ClusterArgs& a,
                                            const GpuCameraUniforms* __restrict__ cam,
                                            const CullLight& L, float sp, MarkBox& m) {
    const int gx = (int)a.grid_x, gy = (int)a.grid_y, gz = (int)a.grid_z;
    const uint32_t lane = hl_lane();
    const float sn = lane < a.grid_z ? a.slice_near[lane] : 0.0f;
    const float sf = lane < a.grid_z ? a.slice_far[lane] : 0.0f;
    int x0 = 0, x1 = gx - 1, y0 = 0, y1 = gy - 1, s0 = 0, s1 = gz - 1;
    if (!__builtin_isinf(L.vpos.w)) {
        const float zc = -L.vpos.z;
        // slices under the largest padding, then again under the largest
        // padding of those slices (paddings grow with distance)
        float rr = L.vpos.w + a.r_pad_column;
        s1 = clampi(lane_lower_bound(sn, gz, zc + rr) + 1, 0, gz - 1);
        rr = L.vpos.w + hl_readlane(sp, (uint32_t)s1);
        s0 = lane_lower_bound(sf, gz, zc - rr) - 1;
        s1 = lane_lower_bound(sn, gz, zc + rr) + 1;
        if (zc > 1e-4f) {
            // A lateral plane through the origin at angle alpha from the view
            // axis passes a centre up to rr / cos(alpha) beyond the column's
            // edge at the centre's own depth; 1/cos(alpha) <= sqrt(1 + 1/p00^2).
            // fast arithmetic: a bound of candidates, widened by a whole cell
            // on each side below, far beyond its rounding
            const float p00 = cam->proj[0], p11 = cam->proj[5];
            const float fx = __builtin_amdgcn_sqrtf(1.0f + hl_f_rcp(p00 * p00));
            const float fy = __builtin_amdgcn_sqrtf(1.0f + hl_f_rcp(p11 * p11));
            const float iz = hl_f_rcp(zc);
            const float nx0 = p00 * (L.vpos.x - rr * fx) * iz, nx1 = p00 * (L.vpos.x + rr * fx) * iz;
            const float ny0 = p11 * (L.vpos.y - rr * fy) * iz, ny1 = p11 * (L.vpos.y + rr * fy) * iz;
            const float fw = (float)a.width, fh = (float)a.height;
            const float px0 = (nx0 + 1.0f) * 0.5f * fw, px1 = (nx1 + 1.0f) * 0.5f * fw;
            const float py0 = (1.0f - ny1) * 0.5f * fh, py1 = (1.0f - ny0) * 0.5f * fh;
            const float icw = hl_f_rcp((float)a.cell_w), ich = hl_f_rcp((float)a.cell_h);
            x0 = (int)fmaxf(floorf(px0 * icw) - 1.0f, -1.0f);
            x1 = (int)fminf(floorf(px1 * icw) + 1.0f, (float)gx);
            y0 = (int)fmaxf(floorf(py0 * ich) - 1.0f, -1.0f);
            y1 = (int)fminf(floorf(py1 * ich) + 1.0f, (float)gy);
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

__device__ __forceinline__ void chunk_mark(const CullLight& L, const ChunkCell& r, uint32_t li,
                                           uint32_t* __restrict__ bits, uint32_t* __restrict__ summary,
                                           uint32_t W, uint32_t SW) {
    if (!r.on || !cell_member(L, r.g, r.bx, r.pad)) return;
    const uint32_t word = li >> 5;
    __hip_atomic_fetch_or(hl_g(&bits[r.cell * W + word]), 1u << (li & 31u), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(hl_g(&summary[r.cell * SW + (word >> 5)]), 1u << (word & 31u), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void mark_chunk(const ClusterArgs& a, const CullLight& L,
                                           const MarkBox& m, uint32_t li,
                                           const ColumnGeom* __restrict__ geom,
                                           const CellBox* __restrict__ boxes,
                                           uint32_t* __restrict__ bits,
                                           uint32_t* __restrict__ summary, uint32_t W,
                                           uint32_t SW, uint32_t chunk) {
    chunk_mark(L, chunk_fetch(a, m, geom, boxes, chunk), li, bits, summary, W, SW);
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
                                           uint32_t* __restrict__ summary, uint32_t W, uint32_t SW,
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
        const uint32_t fit = base >= q.seg ? 0u : min(extra, q.seg - base);
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
    chunk_mark(L, r0, li, bits, summary, W, SW);
    if (nchunk > 1u) chunk_mark(L, r1, li, bits, summary, W, SW);
    for (uint32_t c = 2u; c < qlo; ++c) mark_chunk(a, L, m, li, geom, boxes, bits, summary, W, SW, c);
    for (uint32_t c = qhi; c < nchunk; ++c) mark_chunk(a, L, m, li, geom, boxes, bits, summary, W, SW, c);
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
// the cluster, in ascending order, uniformly over the cluster's 16 lanes (SW
// is uniform over them).  clear: zero the summary words afterwards.
template <class F>
__device__ __forceinline__ void for_nonempty_words(uint32_t* __restrict__ summary_cell, uint32_t SW,
                                                   bool on, bool clear, uint32_t* list, F&& f) {
    const uint32_t sl = threadIdx.x & (kCellLanes - 1u);
    for (uint32_t s0 = 0; s0 < SW; s0 += kCellLanes) {
        const uint32_t j = s0 + sl;
        const uint32_t sw = (on && j < SW) ? summary_cell[j] : 0u;
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
        if (clear && sw) summary_cell[j] = 0u;
    }
}

// number of members of a cluster (before the cap), returned to every lane of
// the cluster's group; lane 0 of the group also stores it in raw[cell]
__device__ __forceinline__ uint32_t count_cell(uint32_t ncell, uint32_t W, uint32_t SW,
                                               const uint32_t* __restrict__ bits,
                                               const uint32_t* __restrict__ summary,
                                               uint32_t* __restrict__ raw, uint32_t cell,
                                               CellSmem& sm) {
    const uint32_t sl = threadIdx.x & (kCellLanes - 1u);
    const bool on = cell < ncell;
    const uint32_t* cb = bits + (size_t)(on ? cell : 0u) * W;
    uint32_t total = 0;
    for_nonempty_words(const_cast<uint32_t*>(summary) + (size_t)(on ? cell : 0u) * SW, SW, on, false,
                       sm.words[threadIdx.x / kCellLanes], [&](const uint32_t* list, uint32_t n) {
                           uint32_t wv[kWordsPerLane];
#pragma unroll
                           for (uint32_t q = 0; q < kWordsPerLane; ++q) {
                               const uint32_t e = kWordsPerLane * sl + q;
                               wv[q] = cb[list[e < n ? e : 0u]];
                           }
#pragma unroll
                           for (uint32_t q = 0; q < kWordsPerLane; ++q)
                               if (kWordsPerLane * sl + q < n) total += (uint32_t)__popc(wv[q]);
                       });
#pragma unroll
    for (uint32_t o = kCellLanes / 2; o >= 1; o >>= 1) total += __shfl_xor(total, o, kCellLanes);
    if (on && sl == 0) raw[cell] = total;
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

// Block per scan tile (256 threads x 4 consecutive clusters): offsets are
// the exclusive scan of the capped counts in cluster order, from the sums of
// the preceding tiles plus the scan within the tile; pool truncation and
// overflow flags (OR-ed into flags_out and, when given, flags_out2).
constexpr uint32_t kScanVPT = kScanTile / 256u;
struct ScanSmem {
    unsigned long long wave[4];
    unsigned long long red[4];
    uint32_t flags;
};

__device__ __forceinline__ void scan_tile(const uint32_t* __restrict__ raw, uint32_t ncell,
                                          uint32_t cap, uint32_t pool_cap,
                                          uint32_t* __restrict__ offsets,
                                          uint32_t* __restrict__ counts,
                                          uint32_t* __restrict__ flags_out,
                                          uint32_t* __restrict__ flags_out2, uint32_t tile,
                                          ScanSmem& sm) {
    const unsigned long long* tile_sums =
        (const unsigned long long*)((const uint8_t*)raw + cluster_tile_sums_offset(ncell));
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    if (t == 0) sm.flags = 0;
    // sums of the preceding tiles
    unsigned long long before = 0;
    for (uint32_t j = t; j < tile; j += 256u) before += tile_sums[j];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) before += __shfl_xor(before, o, 64);
    if (lane == 0) sm.red[wave] = before;
    const uint32_t c0 = tile * kScanTile + kScanVPT * t;
    uint32_t v[kScanVPT];
    if (c0 + kScanVPT <= ncell) {  // raw is 256-B aligned, c0 a multiple of 4
#pragma unroll
        for (uint32_t q = 0; q < kScanVPT; q += 4) {
            const uint4 u = *reinterpret_cast<const uint4*>(raw + c0 + q);
            v[q] = u.x;
            v[q + 1] = u.y;
            v[q + 2] = u.z;
            v[q + 3] = u.w;
        }
    } else {
#pragma unroll
        for (uint32_t q = 0; q < kScanVPT; ++q) v[q] = c0 + q < ncell ? raw[c0 + q] : 0u;
    }
    uint32_t fl = 0;
    unsigned long long local = 0;
#pragma unroll
    for (uint32_t q = 0; q < kScanVPT; ++q) {
        if (cap && v[q] > cap) {
            v[q] = cap;
            fl |= HL_OVERFLOW_CAP;
        }
        local += v[q];
    }
    unsigned long long incl = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long u = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += u;
    }
    if (lane == 63) sm.wave[wave] = incl;
    __syncthreads();
    unsigned long long off = sm.red[0] + sm.red[1] + sm.red[2] + sm.red[3] + (incl - local);
    for (uint32_t w = 0; w < wave; ++w) off += sm.wave[w];
#pragma unroll
    for (uint32_t q = 0; q < kScanVPT; ++q) {
        const uint32_t c = c0 + q;
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
    if (fl) atomicOr(&sm.flags, fl);
    __syncthreads();
    if (t == 0 && sm.flags) {
        atomicOr(flags_out, sm.flags);
        if (flags_out2) atomicOr(flags_out2, sm.flags);
    }
}

// members of a cluster in ascending order up to its (capped, pool-truncated)
// count; clears the cluster's bitmask words for the next call
__device__ __forceinline__ void fill_cell(uint32_t ncell, uint32_t W, uint32_t SW,
                                          uint32_t* __restrict__ bits,
                                          uint32_t* __restrict__ summary,
                                          const uint32_t* __restrict__ offsets,
                                          const uint32_t* __restrict__ counts,
                                          uint32_t* __restrict__ pool, uint32_t cell, CellSmem& sm) {
    const uint32_t sl = threadIdx.x & (kCellLanes - 1u);
    const bool on = cell < ncell;
    const uint32_t cnt = on ? counts[cell] : 0u;
    uint32_t* dst = on ? pool + offsets[cell] : pool;
    uint32_t* cb = bits + (size_t)(on ? cell : 0u) * W;
    uint32_t written = 0;
    for_nonempty_words(summary + (size_t)(on ? cell : 0u) * SW, SW, on, true,
                       sm.words[threadIdx.x / kCellLanes], [&](const uint32_t* list, uint32_t n) {
                           uint32_t wi[kWordsPerLane], wv[kWordsPerLane];
#pragma unroll
                           for (uint32_t q = 0; q < kWordsPerLane; ++q) {
                               const uint32_t e = kWordsPerLane * sl + q;
                               wi[q] = list[e < n ? e : 0u];
                               wv[q] = cb[wi[q]];
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
                               cb[wi[q]] = 0u;
                               for (uint32_t m = wv[q]; m; m &= m - 1u, ++pos)
                                   if (pos < cnt) dst[pos] = 32u * wi[q] + (uint32_t)__builtin_ctz(m);
                           }
                           written += __shfl(incl, kCellLanes - 1u, kCellLanes);
                       });
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ------------------------------------------------------ one view per launch
// The geometry kernel also resets the mark queue and the overflow words (it
// runs before every kernel that reads them).  Its blocks past the cells'
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
                                                      uint32_t* __restrict__ summary, uint32_t W,
                                                      uint32_t SW, MarkQueue q) {
    const uint32_t li = blockIdx.x * 4u + (threadIdx.x >> 6);
    mark_light(a, cam, cl, geom, boxes, bits, summary, W, SW, li, q, li, 0u);
}

// Queued chunks over a fixed grid of wavefronts: nw / kMarkQueues waves per
// sub-queue stride over its valid entries, min(count, seg), then exit.
__global__ __launch_bounds__(256) void k_cluster_mark_tail(ClusterArgs a,
                                                           const CullLight* __restrict__ cl,
                                                           const ColumnGeom* __restrict__ geom,
                                                           const CellBox* __restrict__ boxes,
                                                           uint32_t* __restrict__ bits,
                                                           uint32_t* __restrict__ summary,
                                                           uint32_t W, uint32_t SW, MarkQueue q) {
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
        mark_chunk(a, L, m, li, geom, boxes, bits, summary, W, SW, chunk);
    }
}

__global__ __launch_bounds__(256) void k_cluster_count(uint32_t ncell, uint32_t W, uint32_t SW,
                                                       const uint32_t* __restrict__ bits,
                                                       const uint32_t* __restrict__ summary,
                                                       uint32_t* __restrict__ raw, uint32_t cap) {
    __shared__ TileSumSmem sm;
    __shared__ CellSmem cs;
    const uint32_t cell = (blockIdx.x * blockDim.x + threadIdx.x) / kCellLanes;
    const uint32_t n = count_cell(ncell, W, SW, bits, summary, raw, cell, cs);
    unsigned long long* ts = (unsigned long long*)((uint8_t*)raw + cluster_tile_sums_offset(ncell));
    add_tile_sums(sm, cell < ncell ? ts + cell / kScanTile : nullptr, (cap && n > cap) ? cap : n);
}

__global__ __launch_bounds__(256) void k_cluster_scan(const uint32_t* __restrict__ raw,
                                                      uint32_t ncell, uint32_t cap,
                                                      uint32_t pool_cap,
                                                      uint32_t* __restrict__ offsets,
                                                      uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ flags_out,
                                                      uint32_t* __restrict__ flags_out2) {
    __shared__ ScanSmem sm;
    scan_tile(raw, ncell, cap, pool_cap, offsets, counts, flags_out, flags_out2, blockIdx.x, sm);
}

__global__ __launch_bounds__(256) void k_cluster_fill(uint32_t ncell, uint32_t W, uint32_t SW,
                                                      uint32_t* __restrict__ bits,
                                                      uint32_t* __restrict__ summary,
                                                      const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ counts,
                                                      uint32_t* __restrict__ pool) {
    __shared__ CellSmem cs;
    fill_cell(ncell, W, SW, bits, summary, offsets, counts, pool,
              (blockIdx.x * blockDim.x + threadIdx.x) / kCellLanes, cs);
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

// Cells of every view, then (blocks past geom_blocks) the lights of every
// view gathered from the pool, in one launch.
__global__ __launch_bounds__(256) void k_batch_geom(const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ cell_base,
                                                    uint32_t nviews, uint32_t* __restrict__ qcount,
                                                    const GpuLight* __restrict__ pool,
                                                    const uint32_t* __restrict__ light_base,
                                                    uint32_t geom_blocks) {
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
               v.W, v.SW, g - light_base[vi], q, g, vi);
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
        mark_chunk(v.ca, L, m, li, (const ColumnGeom*)v.geom, view_boxes(v), v.bits, v.summary, v.W,
                   v.SW, chunk);
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
    const uint32_t n = count_cell(v.ncell, v.W, v.SW, v.bits, v.summary, v.raw, c, cs);
    unsigned long long* ts = (unsigned long long*)((uint8_t*)v.raw + cluster_tile_sums_offset(v.ncell));
    const uint32_t cap = v.ca.cap;
    add_tile_sums(sm, c < v.ncell ? ts + c / kScanTile : nullptr, (cap && n > cap) ? cap : n);
}

// one block per scan tile of any view
__global__ __launch_bounds__(256) void k_batch_scan(const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ tile_base,
                                                    uint32_t nviews) {
    __shared__ ScanSmem sm;
    const uint32_t vi = find_view(tile_base, nviews, blockIdx.x);
    const BatchCullView& v = views[vi];
    scan_tile(v.raw, v.ncell, v.ca.cap, v.pool_cap, v.offsets, v.counts, v.flags, nullptr,
              blockIdx.x - tile_base[vi], sm);
}

__global__ __launch_bounds__(256) void k_batch_fill(const BatchCullView* __restrict__ views,
                                                    const uint32_t* __restrict__ cell_base,
                                                    uint32_t nviews) {
    constexpr uint32_t kCellsPerBlock = 256u / kCellLanes;
    const uint32_t vi = find_view(cell_base, nviews, blockIdx.x * kCellsPerBlock);
    const BatchCullView& v = views[vi];
    __shared__ CellSmem cs;
    fill_cell(v.ncell, v.W, v.SW, v.bits, v.summary, v.offsets, v.counts, v.pool,
              blockIdx.x * kCellsPerBlock + threadIdx.x / kCellLanes - cell_base[vi], cs);
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
    k_batch_scan<<<total_tiles, 256, 0, s>>>(views, tile_base, nviews);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_batch_fill<<<cblocks, 256, 0, s>>>(views, cell_base, nviews);
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
                               const GpuLight* lights, HotLight* hot, CullLight* cl, uint32_t* feature_flags,
                               const ClusterScratch& w, uint32_t* offsets,
                               uint32_t* counts, uint32_t* pool, uint32_t pool_cap,
                               uint32_t* flags, uint32_t* flags2) {
    const uint32_t ncol = a.grid_x * a.grid_y;
    const uint32_t ncell = ncol * a.grid_z;
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
                                                                  w.summary, W, SW, w.mq);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        k_cluster_mark_tail<<<tail_blocks(a.num_lights), 256, 0, s>>>(a, cl, geom, boxes, w.bits,
                                                                       w.summary, W, SW, w.mq);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const uint32_t blocks = (ncell + 256u / kCellLanes - 1u) / (256u / kCellLanes);
    k_cluster_count<<<blocks, 256, 0, s>>>(ncell, W, SW, w.bits, w.summary, w.raw, a.cap);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_cluster_scan<<<ntiles, 256, 0, s>>>(w.raw, ncell, a.cap, pool_cap, offsets, counts, flags, flags2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_cluster_fill<<<blocks, 256, 0, s>>>(ncell, W, SW, w.bits, w.summary, offsets, counts, pool);
    return hipGetLastError();
}

}  // namespace hl
