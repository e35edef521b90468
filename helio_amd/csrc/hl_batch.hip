// hl_batch.hip -- batched multi-view frames: one plan holds many views (probe
// cube faces, XR eyes) over a shared light pool.  A frame is a fixed sequence
// of launches that each cover every view:
//   light gather + clustered cull (k_batch_prep/geom/mark/count/scan/fill),
//   one persistent shading launch over a heaviest-first queue of
//   (view, block tile) items (forward, or forward + L2 adjoint per view),
//   per-light gradient write-back for the adjoint views;
// hl_batch_capture records that sequence into a HIP graph once and
// hl_batch_launch replays it.  The autotuner picks, per view class, the
// wavefront tile shape and the cluster grid (cell size and depth slices) by
// timing candidates on the device, optionally caching the choice in a file.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "hl_internal.hpp"

using namespace hl;

namespace {

struct ViewCfg {
    uint32_t tw_log2;  // 3: 8x8, 4: 16x4, 5: 32x2
    uint32_t cell;     // cluster cell size in pixels
    uint32_t gz;       // depth slices
    float ms;
};

struct ViewBufs {
    void* hot = nullptr;
    void* cull = nullptr;
    void* geom = nullptr;
    void* bits = nullptr;
    void* summary = nullptr;
    void* raw = nullptr;
    void* offsets = nullptr;
    void* counts = nullptr;
    void* pool = nullptr;
    void* acc = nullptr;
    uint32_t pool_cap = 0;
};

}  // namespace

struct hl_plan {
    hl_context* ctx = nullptr;
    const GpuLight* pool = nullptr;
    uint32_t pool_size = 0;
    std::vector<hl_view_desc> views;
    std::vector<uint32_t> nlights;
    std::vector<ViewCfg> cfg;
    std::vector<hl_cluster_params> cparams;
    std::vector<ViewBufs> bufs;
    std::vector<BatchCullView> hcull;
    std::vector<BatchShadeView> hshade;
    BatchCullView* d_cull = nullptr;
    BatchShadeView* d_shade = nullptr;
    uint32_t* d_light_base = nullptr;
    uint32_t* d_cell_base = nullptr;
    uint32_t* d_tile_base = nullptr;
    uint2* d_spans = nullptr;  // per group: (first item, view) per view + (item count, 0)
    uint32_t* d_queue = nullptr;
    uint32_t* d_flags = nullptr;
    void* d_markq = nullptr;  // mark_queue_bytes(total_lights)
    // the adjoint views' gradient records: cleared by one launch ahead of the
    // shading, summed into the FP32 outputs by one after it
    GradView* d_grad = nullptr;
    uint32_t num_grad = 0, grad_max_n = 0;
    size_t grad_words = 0;
    uint32_t total_lights = 0, total_cells = 0, total_tiles = 0, num_items = 0;
    // shading items grouped by (G-buffer format set, mode): one persistent
    // launch per non-empty group; group = set * 2 + (adjoint ? 1 : 0)
    uint32_t grp_first[6] = {}, grp_spans[6] = {}, grp_count[6] = {}, grp_grid[6] = {};
    uint32_t grp_view[6] = {};  // the view of a group of one view
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

hl_status plan_alloc(hl_plan* p, void** out, size_t bytes, bool zero = false) {
    hl_context* c = p->ctx;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(c, HL_ENOMEM, "plan allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    if (zero) HL_HIP(c, hipMemset(*out, 0, bytes));
    return HL_OK;
}

void free_bufs(ViewBufs& b) {
    for (void** q : {&b.hot, &b.cull, &b.geom, &b.bits, &b.summary, &b.raw, &b.offsets, &b.counts,
                     &b.pool, &b.acc})
        if (*q) {
            (void)hipFree(*q);
            *q = nullptr;
        }
}

void free_plan_device(hl_plan* p) {
    for (auto& b : p->bufs) free_bufs(b);
    for (void* q : {(void*)p->d_cull, (void*)p->d_shade, (void*)p->d_light_base, (void*)p->d_cell_base,
                    (void*)p->d_tile_base,
                    (void*)p->d_spans, (void*)p->d_queue, (void*)p->d_flags, p->d_markq})
        if (q) (void)hipFree(q);
    p->d_markq = nullptr;
    p->d_cull = nullptr;
    p->d_shade = nullptr;
    p->d_light_base = p->d_cell_base = p->d_tile_base = p->d_queue = p->d_flags = nullptr;
    p->d_spans = nullptr;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

bool is_adjoint(const hl_view_desc& v) { return v.target_d != nullptr; }

// Relative cost of one block tile of a view (for the heaviest-first order).
double block_cost(const hl_plan* p, uint32_t v) {
    const double lights = (double)p->nlights[v];
    return (is_adjoint(p->views[v]) ? 4.0 : 1.0) * (1.0 + log2(1.0 + lights));
}

hl_status enqueue_frame(hl_plan* p, hipStream_t s) {
    hl_context* c = p->ctx;
    const uint32_t nv = (uint32_t)p->views.size();
    HL_HIP(c, launch_grad_views_zero(s, p->d_grad, p->num_grad, p->grad_words));
    // the cull's first kernel clears the views' overflow words; the shading
    // launches leave their claim counters zero for the next frame
    HL_HIP(c, launch_batch_cull(s, p->pool, p->d_cull, p->d_light_base, p->d_cell_base,
                                p->d_tile_base, nv, p->total_lights, p->total_cells, p->total_tiles,
                                mark_queue_at(p->d_markq, p->total_lights)));
    for (int g = 0; g < 6; ++g)
        HL_HIP(c, launch_shade_batch(s, g / 2, (g & 1) ? kModeFusedL2 : kModeForward, p->d_shade,
                                     p->d_spans + p->grp_first[g], p->grp_spans[g], p->grp_count[g],
                                     p->d_queue + (size_t)g * kShadeQueueWords, p->grp_grid[g],
                                     p->grp_spans[g] == 1u ? &p->hshade[p->grp_view[g]].sa : nullptr));
    HL_HIP(c, launch_grad_views_to_f32(s, p->d_grad, p->num_grad, p->grad_max_n));
    return HL_OK;
}

// Builds every device structure of the plan from p->views / p->cfg.
hl_status build_plan(hl_plan* p) {
    hl_context* c = p->ctx;
    const uint32_t nv = (uint32_t)p->views.size();
    p->cparams.assign(nv, hl_cluster_params{});
    p->bufs.assign(nv, ViewBufs{});
    p->hcull.assign(nv, BatchCullView{});
    p->hshade.assign(nv, BatchShadeView{});
    std::vector<uint32_t> light_base(nv + 1, 0), cell_base(nv + 1, 0), tile_base(nv + 1, 0);
    hl_status s;
    if ((s = plan_alloc(p, (void**)&p->d_flags, 4 * (size_t)nv)) != HL_OK) return s;
    for (uint32_t v = 0; v < nv; ++v) {
        const hl_view_desc& d = p->views[v];
        const ViewCfg& cf = p->cfg[v];
        const uint32_t W = d.gbuffer.width, H = d.gbuffer.height;
        hl_cluster_params& cp = p->cparams[v];
        if (d.clusters) {
            cp = *d.clusters;  // checked by hl_batch_plan_create
        } else {
            const uint32_t gx = (W + cf.cell - 1) / cf.cell, gy = (H + cf.cell - 1) / cf.cell;
            hl_status st = hl_cluster_params_init(&d.camera, W, H, gx, gy, cf.gz, 0, &cp);
            if (st != HL_OK) return fail(c, st, "view %u: camera is not a standard-Z perspective", v);
        }
        // The grid bounds of check_cluster_params (at most 65535 columns and
        // 2^26 cells) keep the packed fields of the mark queue in range: box
        // extents and slice fields in 16 bits, chunk indices (ncell / 64 <
        // 2^20) in the 24 bits beside the 8-bit view index.
        {
            const hl_status st = check_cluster_params(c, &cp);
            if (st != HL_OK) return st;
        }
        static_assert(HL_BATCH_MAX_VIEWS < 256u, "view index of a mark queue entry: 8 bits");
        static_assert((1u << 26) / 64u < (1u << 24), "chunk index of a mark queue entry: 24 bits");
        const uint32_t n = p->nlights[v];
        const uint32_t ncol = cp.grid_x * cp.grid_y, ncell = ncol * cp.grid_z;
        ViewBufs& b = p->bufs[v];
        if ((s = plan_alloc(p, &b.hot, sizeof(HotLight) * (size_t)(n ? n : 1))) != HL_OK) return s;
        if ((s = plan_alloc(p, &b.cull, sizeof(CullLight) * (size_t)(n ? n : 1))) != HL_OK) return s;
        if ((s = plan_alloc(p, &b.geom, cluster_geom_bytes(ncol, ncell))) != HL_OK) return s;
        if ((s = plan_alloc(p, &b.bits, 4 * cluster_bits_words(ncell, n), true)) != HL_OK) return s;
        if ((s = plan_alloc(p, &b.summary, 4 * cluster_summary_words(ncell, n), true)) != HL_OK) return s;
        if ((s = plan_alloc(p, &b.raw, cluster_raw_bytes(ncell))) != HL_OK) return s;
        if ((s = plan_alloc(p, &b.offsets, 4 * (size_t)ncell)) != HL_OK) return s;
        if ((s = plan_alloc(p, &b.counts, 4 * (size_t)ncell)) != HL_OK) return s;
        // provisional pool; sized exactly after a first cull below
        b.pool_cap = (uint32_t)std::min<uint64_t>((uint64_t)ncell * std::min<uint32_t>(n, 32u) + 4096u,
                                                  0x7fffffffu);
        if ((s = plan_alloc(p, &b.pool, 4 * (size_t)b.pool_cap)) != HL_OK) return s;
        if (is_adjoint(d) && (s = plan_alloc(p, &b.acc, sizeof(double) * grad_accum_doubles(n))) != HL_OK)
            return s;
        BatchCullView& cv = p->hcull[v];
        cv.ca = make_cluster_args(&cp, n);
        cv.cam = d.camera_d;
        cv.subset = d.light_indices_d;
        cv.hot = (HotLight*)b.hot;
        cv.cull = (CullLight*)b.cull;
        cv.geom = b.geom;
        cv.bits = (uint32_t*)b.bits;
        cv.summary = (uint32_t*)b.summary;
        cv.raw = (uint32_t*)b.raw;
        cv.offsets = (uint32_t*)b.offsets;
        cv.counts = (uint32_t*)b.counts;
        cv.pool = (uint32_t*)b.pool;
        cv.flags = p->d_flags + v;
        cv.pool_cap = b.pool_cap;
        cv.W = (n + 31u) / 32u;
        cv.SW = (cv.W + 31u) / 32u;
        cv.ncell = ncell;
        if ((uint64_t)light_base[v] + n >= (1ull << 28) ||
            (uint64_t)cell_base[v] + ncell + kBatchCellAlign >= (1ull << 31))
            return fail(c, HL_ECAPACITY, "plan too large: %u views up to view %u exceed 2^28 lights or 2^31 cells",
                        nv, v);
        light_base[v + 1] = light_base[v] + n;
        cell_base[v + 1] = cell_base[v] + (ncell + kBatchCellAlign - 1u) / kBatchCellAlign * kBatchCellAlign;
        tile_base[v + 1] = tile_base[v] + cluster_scan_tiles(ncell);
    }
    p->total_lights = light_base[nv];
    p->total_cells = cell_base[nv];
    p->total_tiles = tile_base[nv];
    if ((s = plan_alloc(p, (void**)&p->d_cull, sizeof(BatchCullView) * nv)) != HL_OK) return s;
    if ((s = plan_alloc(p, (void**)&p->d_light_base, 4 * (size_t)(nv + 1))) != HL_OK) return s;
    if ((s = plan_alloc(p, (void**)&p->d_cell_base, 4 * (size_t)(nv + 1))) != HL_OK) return s;
    if ((s = plan_alloc(p, &p->d_markq, mark_queue_bytes(p->total_lights))) != HL_OK) return s;
    if ((s = plan_alloc(p, (void**)&p->d_tile_base, 4 * (size_t)(nv + 1))) != HL_OK) return s;
    HL_HIP(c, hipMemcpy(p->d_tile_base, tile_base.data(), 4 * (nv + 1), hipMemcpyHostToDevice));
    HL_HIP(c, hipMemcpy(p->d_light_base, light_base.data(), 4 * (nv + 1), hipMemcpyHostToDevice));
    HL_HIP(c, hipMemcpy(p->d_cell_base, cell_base.data(), 4 * (nv + 1), hipMemcpyHostToDevice));
    // size each pool from a first cull: exact totals + headroom
    for (int pass = 0; pass < 2; ++pass) {
        HL_HIP(c, hipMemcpy(p->d_cull, p->hcull.data(), sizeof(BatchCullView) * nv, hipMemcpyHostToDevice));
        HL_HIP(c, launch_fill_u32(c->stream, p->d_flags, 0, nv));
        HL_HIP(c, launch_batch_cull(c->stream, p->pool, p->d_cull, p->d_light_base, p->d_cell_base,
                                    p->d_tile_base, nv, p->total_lights, p->total_cells,
                                    p->total_tiles, mark_queue_at(p->d_markq, p->total_lights)));
        {
            const hipError_t e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess)
                return fail(c, HL_EHIP, "plan cull pass %d (%u views, %u lights, %u clusters): %s", pass, nv,
                            p->total_lights, p->total_cells, hipGetErrorString(e));
        }
        bool grew = false;
        for (uint32_t v = 0; v < nv; ++v) {
            std::vector<uint32_t> raw(p->hcull[v].ncell);
            HL_HIP(c, hipMemcpy(raw.data(), p->bufs[v].raw, 4 * raw.size(), hipMemcpyDeviceToHost));
            uint64_t total = 0;
            for (uint32_t r : raw) total += r;
            const uint64_t want = total + total / 2 + 4096;
            if (want > 0x7fffffffull) return fail(c, HL_ECAPACITY, "view %u needs %llu list entries", v,
                                                  (unsigned long long)total);
            if (total > p->bufs[v].pool_cap || pass == 0) {
                (void)hipFree(p->bufs[v].pool);
                p->bufs[v].pool = nullptr;
                p->bufs[v].pool_cap = (uint32_t)want;
                if ((s = plan_alloc(p, &p->bufs[v].pool, 4 * want)) != HL_OK) return s;
                p->hcull[v].pool = (uint32_t*)p->bufs[v].pool;
                p->hcull[v].pool_cap = (uint32_t)want;
                grew = true;
            }
        }
        if (!grew) break;
    }
    HL_HIP(c, hipMemcpy(p->d_cull, p->hcull.data(), sizeof(BatchCullView) * nv, hipMemcpyHostToDevice));
    // shading records
    std::vector<std::pair<double, uint32_t>> order;
    std::vector<int> vset(nv, 0);
    for (uint32_t v = 0; v < nv; ++v) {
        const hl_view_desc& d = p->views[v];
        hl_light_grid grid{};
        grid.mode = HL_GRID_CLUSTERS;
        grid.lists_d = (const uint32_t*)p->bufs[v].pool;
        grid.counts_d = (const uint32_t*)p->bufs[v].counts;
        grid.offsets_d = (const uint32_t*)p->bufs[v].offsets;
        grid.clusters = &p->cparams[v];
        grid.lists_capacity = p->bufs[v].pool_cap;
        int set = 0;
        BatchShadeView& sv = p->hshade[v];
        hl_status st = build_shade_args(c, d.camera_d, &d.globals, &d.gbuffer, p->nlights[v], &grid,
                                        &sv.sa, &set, is_adjoint(d));
        if (st != HL_OK) return st;
        vset[v] = set;
        set_tile_shape(sv.sa, p->cfg[v].tw_log2);
        sv.sa.lights = (const HotLight*)p->bufs[v].hot;
        sv.sa.out = d.out_d;
        sv.sa.out_pitch = d.out_pitch_bytes;
        sv.mode = kModeForward;
        if (is_adjoint(d)) {
            sv.mode = kModeFusedL2;
            sv.sa.dout = d.target_d;
            sv.sa.dout_pitch = d.target_pitch_bytes;
            sv.sa.loss = d.loss_d;
            sv.sa.d_pixel = d.d_pixel_d;
            sv.sa.d_light = (double*)p->bufs[v].acc;
        }
        order.push_back({block_cost(p, v), v});
    }
    if ((s = plan_alloc(p, (void**)&p->d_shade, sizeof(BatchShadeView) * nv)) != HL_OK) return s;
    HL_HIP(c, hipMemcpy(p->d_shade, p->hshade.data(), sizeof(BatchShadeView) * nv, hipMemcpyHostToDevice));
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<double, uint32_t>& a, const std::pair<double, uint32_t>& b) {
                         return a.first > b.first;
                     });
    // Items of a group: the block tiles of its views, view after view in the
    // order above.  A group is described by spans (first item, view), one per
    // view plus a closing (item count, 0), so an item's view and tile follow
    // from its index.
    std::vector<uint2> spans;
    int cus = 0;
    HL_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    uint64_t all_items = 0;
    for (int g = 0; g < 6; ++g) {
        p->grp_first[g] = (uint32_t)spans.size();
        uint64_t n = 0;
        for (const auto& o : order) {
            const uint32_t v = o.second;
            if (vset[v] * 2 + (is_adjoint(p->views[v]) ? 1 : 0) != g) continue;
            spans.push_back(make_uint2((uint32_t)n, v));
            p->grp_view[g] = v;
            n += shade_items(is_adjoint(p->views[v]) ? kModeFusedL2 : kModeForward, p->hshade[v].sa.num_blocks);
        }
        if (n >= 0x80000000ull) return fail(c, HL_ECAPACITY, "%llu block tiles in one launch", (unsigned long long)n);
        p->grp_spans[g] = (uint32_t)spans.size() - p->grp_first[g];
        spans.push_back(make_uint2((uint32_t)n, 0u));
        p->grp_count[g] = (uint32_t)n;
        all_items += n;
        // the blocks the kernel keeps resident, fewer when there are fewer items
        p->grp_grid[g] = std::min((uint32_t)std::max(1, cus) * shade_batch_blocks_per_cu(g / 2, (g & 1) ? kModeFusedL2
                                                                                                    : kModeForward),
                                  std::max(1u, p->grp_count[g]));
    }
    p->num_items = (uint32_t)all_items;
    if ((s = plan_alloc(p, (void**)&p->d_spans, sizeof(uint2) * spans.size())) != HL_OK) return s;
    HL_HIP(c, hipMemcpy(p->d_spans, spans.data(), sizeof(uint2) * spans.size(), hipMemcpyHostToDevice));
    if ((s = plan_alloc(p, (void**)&p->d_queue, 4 * 6 * (size_t)kShadeQueueWords, true)) != HL_OK) return s;
    std::vector<GradView> gv;
    p->grad_words = 0;
    p->grad_max_n = 0;
    for (uint32_t v = 0; v < nv; ++v) {
        const hl_view_desc& d = p->views[v];
        if (!is_adjoint(d)) continue;
        const GradView g{(double*)p->bufs[v].acc, d.loss_d, d.d_light_pos_d, d.d_light_intensity_d,
                         d.light_indices_d, p->nlights[v], p->pool_size};
        gv.push_back(g);
        p->grad_words = std::max({p->grad_words, 2 * grad_accum_doubles(g.n), 3 * (size_t)g.pool});
        p->grad_max_n = std::max(p->grad_max_n, g.n);
    }
    p->num_grad = (uint32_t)gv.size();
    if (!gv.empty()) {
        if ((s = plan_alloc(p, (void**)&p->d_grad, sizeof(GradView) * gv.size())) != HL_OK) return s;
        HL_HIP(c, hipMemcpy(p->d_grad, gv.data(), sizeof(GradView) * gv.size(), hipMemcpyHostToDevice));
    }
    return HL_OK;
}

// Views of one class share a tuned configuration.  A view with a fixed
// cluster grid keys on that grid too (only its wave tile is tuned).
std::string class_key(const hl_view_desc& d, uint32_t n) {
    char buf[160];
    int bucket = n ? (int)floor(log2((double)n)) : -1;
    if (d.clusters)
        snprintf(buf, sizeof buf, "gfx950 %u %u %d %d fixed-%ux%ux%u-%u", d.gbuffer.width,
                 d.gbuffer.height, is_adjoint(d) ? 1 : 0, bucket, d.clusters->grid_x,
                 d.clusters->grid_y, d.clusters->grid_z, d.clusters->cap);
    else
        snprintf(buf, sizeof buf, "gfx950 %u %u %d %d", d.gbuffer.width, d.gbuffer.height,
                 is_adjoint(d) ? 1 : 0, bucket);
    return buf;
}

std::vector<ViewCfg> candidates(bool fixed_grid) {
    std::vector<ViewCfg> out;
    // default first, then the rest of the grid
    out.push_back({3, 64, 24, 0.0f});
    if (fixed_grid) {
        out.push_back({4, 64, 24, 0.0f});
        return out;
    }
    for (uint32_t tw : {3u, 4u})
        for (uint32_t cell : {32u, 64u, 128u})
            for (uint32_t gz : {16u, 24u, 32u, 48u})
                if (!(tw == 3 && cell == 64 && gz == 24)) out.push_back({tw, cell, gz, 0.0f});
    return out;
}

// Median time of a one-view plan with config cf (cull + shade + write-back).
hl_status time_config(hl_context* c, const GpuLight* pool, uint32_t pool_size,
                      const hl_view_desc& d, uint32_t n, const ViewCfg& cf, float* ms) {
    hl_plan tp;
    tp.ctx = c;
    tp.pool = pool;
    tp.pool_size = pool_size;
    tp.views = {d};
    tp.nlights = {n};
    tp.cfg = {cf};
    hl_status s = build_plan(&tp);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<float> t;
    hipError_t e = hipSuccess;
    if (s == HL_OK) {
        if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
            s = fail(c, HL_EHIP, "autotune: hipEventCreate: %s", hipGetErrorString(e));
        for (int r = 0; r < 6 && s == HL_OK; ++r) {
            if ((e = hipEventRecord(e0, c->stream)) != hipSuccess) break;
            if ((s = enqueue_frame(&tp, c->stream)) != HL_OK) break;
            if ((e = hipEventRecord(e1, c->stream)) != hipSuccess) break;
            // a fault of this frame's kernels surfaces at the stream synchronisation
            if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) break;
            float x = 0.0f;
            if ((e = hipEventElapsedTime(&x, e0, e1)) != hipSuccess) break;
            if (r > 0) t.push_back(x);
        }
        if (s == HL_OK && e != hipSuccess)
            s = fail(c, HL_EHIP, "autotune frame (view %ux%u, tile %u, cell %u, slices %u): %s",
                     d.gbuffer.width, d.gbuffer.height, 1u << cf.tw_log2, cf.cell, cf.gz,
                     hipGetErrorString(e));
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess && s == HL_OK)
        s = fail(c, HL_EHIP, "autotune: hipStreamSynchronize: %s", hipGetErrorString(e));
    free_plan_device(&tp);
    if (s != HL_OK) return s;
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return HL_OK;
}

// One line per view class: "<class key> | <tile log2> <cell px> <slices> <ms>".
void load_cache(const char* path, std::map<std::string, ViewCfg>& cache) {
    if (!path) return;
    FILE* f = fopen(path, "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        const char* bar = strchr(line, '|');
        if (!bar || bar == line) continue;
        unsigned tw, cell, gz;
        float ms;
        if (sscanf(bar + 1, "%u %u %u %f", &tw, &cell, &gz, &ms) != 4 || tw < 3 || tw > 5 || !cell || !gz ||
            gz > HL_MAX_SLICES)
            continue;
        std::string key(line, (size_t)(bar - line));
        while (!key.empty() && key.back() == ' ') key.pop_back();
        cache[key] = ViewCfg{tw, cell, gz, ms};
    }
    fclose(f);
}

void save_cache(const char* path, const std::map<std::string, ViewCfg>& cache) {
    if (!path) return;
    FILE* f = fopen(path, "w");
    if (!f) return;
    for (const auto& kv : cache)
        fprintf(f, "%s | %u %u %u %.6f\n", kv.first.c_str(), kv.second.tw_log2, kv.second.cell,
                kv.second.gz, kv.second.ms);
    fclose(f);
}

}  // namespace

extern "C" {

HL_API hl_status hl_batch_plan_create(hl_context* c, const GpuLight* pool_d, uint32_t pool_size,
                                      const hl_view_desc* views, uint32_t num_views,
                                      const hl_autotune_opts* opts, hl_plan** out) {
    if (!c) return HL_EINVAL;
    if (!out) return fail(c, HL_EINVAL, "plan output is null");
    *out = nullptr;
    if (!views || num_views == 0) return fail(c, HL_EINVAL, "no views");
    if (num_views > HL_BATCH_MAX_VIEWS) return fail(c, HL_EINVAL, "at most %u views per plan", HL_BATCH_MAX_VIEWS);
    if (pool_size && !pool_d) return fail(c, HL_EINVAL, "light pool is null");
    HL_HIP(c, hipSetDevice(c->device));
    hl_plan* p = new hl_plan();
    p->ctx = c;
    p->pool = pool_d;
    p->pool_size = pool_size;
    p->views.assign(views, views + num_views);
    for (uint32_t v = 0; v < num_views; ++v) {
        const hl_view_desc& d = views[v];
        const uint32_t n = d.light_indices_d ? d.num_lights : pool_size;
        if (d.light_indices_d && d.num_lights > pool_size) {
            delete p;
            return fail(c, HL_EINVAL, "view %u: %u lights from a pool of %u", v, d.num_lights, pool_size);
        }
        if (!d.camera_d || !d.out_d) {
            delete p;
            return fail(c, HL_EINVAL, "view %u: camera or output is null", v);
        }
        if (d.light_indices_d && d.num_lights) {
            // the kernels index the pool through the subset: check it once here
            std::vector<uint32_t> idx(d.num_lights);
            hipError_t e = hipMemcpy(idx.data(), d.light_indices_d, 4 * idx.size(), hipMemcpyDeviceToHost);
            if (e != hipSuccess) {
                delete p;
                return hip_fail(c, e, "reading a view's light subset");
            }
            for (size_t i = 0; i < idx.size(); ++i)
                if (idx[i] >= pool_size || (i && idx[i] <= idx[i - 1])) {
                    delete p;
                    return fail(c, HL_EINVAL, "view %u: light subset must be ascending pool indices", v);
                }
        }
        if (d.out_pitch_bytes < 16u * d.gbuffer.width || d.out_pitch_bytes % 16) {
            delete p;
            return fail(c, HL_EINVAL, "view %u: output pitch invalid", v);
        }
        if (d.clusters) {
            hl_status st = check_cluster_params(c, d.clusters);
            if (st == HL_OK && (d.clusters->width != d.gbuffer.width || d.clusters->height != d.gbuffer.height))
                st = fail(c, HL_EINVAL, "view %u: cluster grid is for %ux%u, view is %ux%u", v,
                          d.clusters->width, d.clusters->height, d.gbuffer.width, d.gbuffer.height);
            if (st != HL_OK) {
                delete p;
                return st;
            }
        }
        if (is_adjoint(d) && (!d.loss_d || !d.d_pixel_d || !d.d_light_pos_d || !d.d_light_intensity_d ||
                              d.target_pitch_bytes < 16u * d.gbuffer.width || d.target_pitch_bytes % 16)) {
            delete p;
            return fail(c, HL_EINVAL, "view %u: adjoint outputs or target pitch invalid", v);
        }
        p->nlights.push_back(n);
    }
    // configuration per view
    p->cfg.assign(num_views, candidates(false)[0]);
    if (opts && opts->enable) {
        std::map<std::string, ViewCfg> cache;
        load_cache(opts->cache_path, cache);
        bool dirty = false;
        for (uint32_t v = 0; v < num_views; ++v) {
            const std::string key = class_key(views[v], p->nlights[v]);
            auto it = cache.find(key);
            if (it == cache.end()) {
                const std::vector<ViewCfg> cands = candidates(views[v].clusters != nullptr);
                const size_t trials =
                    opts->max_trials ? std::min<size_t>(opts->max_trials, cands.size()) : cands.size();
                ViewCfg best = cands[0];
                best.ms = 1e30f;
                for (size_t k = 0; k < trials; ++k) {
                    float ms = 0.0f;
                    hl_status s = time_config(c, pool_d, pool_size, views[v], p->nlights[v], cands[k], &ms);
                    if (s != HL_OK) {
                        delete p;
                        return s;
                    }
                    if (ms < best.ms) {
                        best = cands[k];
                        best.ms = ms;
                    }
                }
                it = cache.emplace(key, best).first;
                dirty = true;
            }
            p->cfg[v] = it->second;
        }
        if (dirty) save_cache(opts->cache_path, cache);
    }
    hl_status s = build_plan(p);
    if (s != HL_OK) {
        (void)hipStreamSynchronize(c->stream);
        free_plan_device(p);
        delete p;
        return s;
    }
    *out = p;
    return HL_OK;
}

HL_API hl_status hl_batch_capture(hl_plan* p) {
    if (!p) return HL_EINVAL;
    hl_context* c = p->ctx;
    HL_HIP(c, hipSetDevice(c->device));
    HL_HIP(c, hipStreamSynchronize(c->stream));
    if (p->exec) {
        (void)hipGraphExecDestroy(p->exec);
        p->exec = nullptr;
    }
    if (p->graph) {
        (void)hipGraphDestroy(p->graph);
        p->graph = nullptr;
    }
    if (!p->cap_stream) HL_HIP(c, hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
    HL_HIP(c, hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeRelaxed));
    hl_status s = enqueue_frame(p, p->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
    if (s != HL_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    p->graph = g;
    HL_HIP(c, hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return HL_OK;
}

HL_API hl_status hl_batch_launch(hl_plan* p) {
    if (!p) return HL_EINVAL;
    hl_context* c = p->ctx;
    if (p->exec) {
        HL_HIP(c, hipGraphLaunch(p->exec, c->stream));
        return HL_OK;
    }
    return enqueue_frame(p, c->stream);
}

HL_API hl_status hl_batch_get_config(const hl_plan* p, uint32_t view, hl_view_config* cfg) {
    if (!p || !cfg || view >= p->views.size()) return HL_EINVAL;
    const ViewCfg& cf = p->cfg[view];
    cfg->wave_tile_w = 1u << cf.tw_log2;
    cfg->wave_tile_h = 64u >> cf.tw_log2;
    cfg->grid_x = p->cparams[view].grid_x;
    cfg->grid_y = p->cparams[view].grid_y;
    cfg->grid_z = p->cparams[view].grid_z;
    cfg->est_ms = cf.ms;
    return HL_OK;
}

HL_API hl_status hl_plan_destroy(hl_plan* p) {
    if (!p) return HL_EINVAL;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    if (p->cap_stream) {
        (void)hipStreamSynchronize(p->cap_stream);
        (void)hipStreamDestroy(p->cap_stream);
    }
    free_plan_device(p);
    delete p;
    return HL_OK;
}

}  // extern "C"
