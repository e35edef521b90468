// hl_api.hip -- the C ABI of libhelio_hip.so: context, scratch memory,
// argument validation and the entry points declared in helio_hip/helio_hip.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "hl_internal.hpp"

using namespace hl;

struct hl_event {
    hipEvent_t ev;
};

namespace hl {

hl_status fail(hl_context* c, hl_status s, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return s;
}

hl_status hip_fail(hl_context* c, hipError_t e, const char* what) {
    return fail(c, HL_EHIP, "%s: %s", what, hipGetErrorString(e));
}

hl_status ensure(hl_context* c, DevBuf& b, size_t bytes, bool zero) {
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return HL_OK;
    HL_HIP(c, hipSetDevice(c->device));
    if (b.p) {
        HL_HIP(c, hipStreamSynchronize(c->stream));
        HL_HIP(c, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, HL_ENOMEM, "scratch allocation of %zu bytes failed: %s", want,
                    hipGetErrorString(e));
    }
    b.bytes = want;
    if (zero) {
        HL_HIP(c, hipMemsetAsync(b.p, 0, want, c->stream));
        HL_HIP(c, hipStreamSynchronize(c->stream));
    }
    return HL_OK;
}

uint32_t texel_bytes(uint32_t fmt) {
    switch (fmt) {
        case HL_FMT_RGBA8_UNORM: return 4;
        case HL_FMT_RGBA16_FLOAT: return 8;
        case HL_FMT_RGBA32_FLOAT: return 16;
        case HL_FMT_R32_FLOAT: return 4;
        case HL_FMT_RG16_FLOAT: return 4;
        default: return 0;
    }
}

// A texture of w x h texels; the kernels address its texels with 32-bit
// byte offsets, so its rows must end below 4 GiB from the base.
hl_status check_tex(hl_context* c, const hl_texture& t, uint32_t w, uint32_t h, const char* name) {
    const uint32_t tb = texel_bytes(t.format);
    if (!t.data_d) return fail(c, HL_EINVAL, "gbuffer.%s: null data pointer", name);
    if (tb == 0) return fail(c, HL_EINVAL, "gbuffer.%s: unknown format %u", name, t.format);
    if ((uint64_t)t.pitch_bytes < (uint64_t)w * tb || t.pitch_bytes % tb != 0)
        return fail(c, HL_EINVAL, "gbuffer.%s: pitch %u too small or not a multiple of %u", name,
                    t.pitch_bytes, tb);
    if ((uint64_t)t.pitch_bytes * h > 0xFFFFFFFFull)
        return fail(c, HL_EUNSUPPORTED, "gbuffer.%s: %u rows of %u bytes exceed 4 GiB", name, h, t.pitch_bytes);
    if (((uintptr_t)t.data_d) % tb != 0)
        return fail(c, HL_EINVAL, "gbuffer.%s: data not aligned to %u bytes", name, tb);
    return HL_OK;
}

hl_status color_set(hl_context* c, const hl_gbuffer* g, int* set);

hl_status check_gbuffer(hl_context* c, const hl_gbuffer* g, int* set) {
    if (!g) return fail(c, HL_EINVAL, "gbuffer is null");
    if (g->width == 0 || g->height == 0) return fail(c, HL_EINVAL, "gbuffer has zero size");
    if (g->width > 32768 || g->height > 32768)
        return fail(c, HL_EINVAL, "gbuffer %ux%u exceeds 32768", g->width, g->height);
    hl_status s;
    if ((s = check_tex(c, g->albedo, g->width, g->height, "albedo")) != HL_OK) return s;
    if ((s = check_tex(c, g->normal, g->width, g->height, "normal")) != HL_OK) return s;
    if ((s = check_tex(c, g->orm, g->width, g->height, "orm")) != HL_OK) return s;
    if ((s = check_tex(c, g->emissive, g->width, g->height, "emissive")) != HL_OK) return s;
    if ((s = check_tex(c, g->depth, g->width, g->height, "depth")) != HL_OK) return s;
    if (g->depth.format != HL_FMT_R32_FLOAT)
        return fail(c, HL_EINVAL, "gbuffer.depth must be R32_FLOAT (Depth32Float)");
    if (g->lightmap_uv.data_d)
        return fail(c, HL_EUNSUPPORTED,
                    "gbuffer.lightmap_uv must be NULL: baked lightmaps are outside the hot path");
    return color_set(c, g, set);
}

// The format set of the four colour targets.
hl_status color_set(hl_context* c, const hl_gbuffer* g, int* set) {
    const uint32_t fa = g->albedo.format, fo = g->orm.format, fn = g->normal.format,
                   fe = g->emissive.format;
    if (fa != fo || fn != fe)
        return fail(c, HL_EUNSUPPORTED, "albedo/orm and normal/emissive must share formats");
    if (fa == HL_FMT_RGBA8_UNORM && fn == HL_FMT_RGBA16_FLOAT) *set = kSetRef;
    else if (fa == HL_FMT_RGBA16_FLOAT && fn == HL_FMT_RGBA16_FLOAT) *set = kSet16;
    else if (fa == HL_FMT_RGBA32_FLOAT && fn == HL_FMT_RGBA32_FLOAT) *set = kSet32;
    else return fail(c, HL_EUNSUPPORTED, "unsupported G-buffer format set (%u,%u)", fa, fn);
    return HL_OK;
}

bool debug_supported(uint32_t m) {
    return m == 0 || m == 1 || m == 2 || m == 4 || m == 5 || m == 20 || m == 21;
}

hl_status check_cluster_params(hl_context* c, const hl_cluster_params* p) {
    if (!p) return fail(c, HL_EINVAL, "cluster params are null");
    if (p->grid_x == 0 || p->grid_y == 0 || p->grid_z == 0 || p->grid_z > HL_MAX_SLICES)
        return fail(c, HL_EINVAL, "cluster grid %ux%ux%u invalid (1..%d slices)", p->grid_x,
                    p->grid_y, p->grid_z, HL_MAX_SLICES);
    if (p->cell_w == 0 || p->cell_h == 0 || p->width == 0 || p->height == 0)
        return fail(c, HL_EINVAL, "cluster cell or screen size is zero");
    if ((uint64_t)p->grid_x * p->cell_w < p->width || (uint64_t)p->grid_y * p->cell_h < p->height)
        return fail(c, HL_EINVAL, "cluster grid does not cover the screen");
    if ((uint64_t)p->grid_x * p->grid_y * p->grid_z > (1ull << 26) ||
        (uint64_t)p->grid_x * p->grid_y > 65535u)
        return fail(c, HL_EINVAL, "cluster grid too large (at most 65535 columns, 2^26 cells)");
    // the shading kernels search the slice thresholds as a sorted table
    for (uint32_t k = 1; k < p->grid_z; ++k)
        if (!(p->depth_split[k] >= p->depth_split[k - 1]))
            return fail(c, HL_EINVAL, "cluster depth_split is not non-decreasing at slice %u", k);
    // the cull bounds a light's slices by searching slice_near and slice_far
    // as sorted tables, under the padding of the farthest slice in reach
    for (uint32_t k = 0; k < p->grid_z; ++k) {
        if (!std::isfinite(p->slice_near[k]) || !std::isfinite(p->slice_far[k]) ||
            !std::isfinite(p->slice_r_pad[k]) || !(p->slice_r_pad[k] >= 0.0f))
            return fail(c, HL_EINVAL, "cluster slice %u: slice_near, slice_far and slice_r_pad must be finite, "
                        "the padding non-negative", k);
        if (k > 0 && !(p->slice_near[k] >= p->slice_near[k - 1] && p->slice_far[k] >= p->slice_far[k - 1] &&
                       p->slice_r_pad[k] >= p->slice_r_pad[k - 1]))
            return fail(c, HL_EINVAL, "cluster slice_near, slice_far and slice_r_pad must be non-decreasing "
                        "(slice %u)", k);
    }
    return HL_OK;
}

ClusterArgs make_cluster_args(const hl_cluster_params* p, uint32_t num_lights) {
    ClusterArgs a;
    memset(&a, 0, sizeof a);
    a.grid_x = p->grid_x;
    a.grid_y = p->grid_y;
    a.grid_z = p->grid_z;
    a.cell_w = p->cell_w;
    a.cell_h = p->cell_h;
    a.width = p->width;
    a.height = p->height;
    a.cap = p->cap;
    a.num_lights = num_lights;
    a.r_pad_column = p->r_pad_column;
    memcpy(a.slice_near, p->slice_near, sizeof a.slice_near);
    memcpy(a.slice_far, p->slice_far, sizeof a.slice_far);
    memcpy(a.slice_r_pad, p->slice_r_pad, sizeof a.slice_r_pad);
    return a;
}

uint32_t* ctx_words(hl_context* c) { return (uint32_t*)c->words.p; }

// Validation and shading arguments of one view (no device work; a->lights is
// left for the caller).
hl_status build_shade_args(hl_context* c, const GpuCameraUniforms* camera_d,
                           const DeferredGlobals* g, const hl_gbuffer* gb, uint32_t num_lights,
                           const hl_light_grid* grid, ShadeArgs* a, int* set, bool need_debug0) {
    if (!c) return HL_EINVAL;
    if (!camera_d) return fail(c, HL_EINVAL, "camera is null");
    if (!g) return fail(c, HL_EINVAL, "globals are null");
    hl_status s = check_gbuffer(c, gb, set);
    if (s != HL_OK) return s;
    if (!debug_supported(g->debug_mode))
        return fail(c, HL_EUNSUPPORTED, "debug_mode %u needs shadow data outside the hot path",
                    g->debug_mode);
    if (need_debug0 && g->debug_mode != 0)
        return fail(c, HL_EUNSUPPORTED, "the adjoint is defined for debug_mode 0 only");
    if (!grid) return fail(c, HL_EINVAL, "light grid is null");
    memset(a, 0, sizeof *a);
    const uint32_t W = gb->width, H = gb->height;
    if (grid->mode == HL_GRID_TILES) {
        const uint32_t tx = (W + 15) / 16, ty = (H + 15) / 16;
        if (!grid->lists_d || !grid->counts_d) return fail(c, HL_EINVAL, "tile lists/counts are null");
        if (grid->num_tiles_x < tx || grid->num_tiles_y < ty)
            return fail(c, HL_EINVAL, "tile grid %ux%u does not cover %ux%u px", grid->num_tiles_x,
                        grid->num_tiles_y, W, H);
        if (g->num_tiles_x != grid->num_tiles_x)
            return fail(c, HL_EINVAL, "globals.num_tiles_x (%u) != grid.num_tiles_x (%u)",
                        g->num_tiles_x, grid->num_tiles_x);
        a->grid_mode = HL_GRID_TILES;
        a->tile_stride_x = g->num_tiles_x;
        uint64_t cap = (uint64_t)grid->num_tiles_x * grid->num_tiles_y * HL_MAX_LIGHTS_PER_TILE;
        if (grid->lists_capacity && grid->lists_capacity < cap) cap = grid->lists_capacity;
        a->lists_capacity = cap > 0xffffffffull ? 0xffffffffu : (uint32_t)cap;
        a->grid_z = 1;
    } else if (grid->mode == HL_GRID_CLUSTERS) {
        const hl_cluster_params* p = grid->clusters;
        if ((s = check_cluster_params(c, p)) != HL_OK) return s;
        if (p->width != W || p->height != H)
            return fail(c, HL_EINVAL, "cluster params are for %ux%u, G-buffer is %ux%u", p->width,
                        p->height, W, H);
        if (!grid->lists_d || !grid->counts_d || !grid->offsets_d)
            return fail(c, HL_EINVAL, "cluster offsets/counts/pool are null");
        a->grid_mode = HL_GRID_CLUSTERS;
        a->grid_x = p->grid_x;
        a->grid_y = p->grid_y;
        a->grid_z = p->grid_z;
        a->cell_w = p->cell_w;
        a->cell_h = p->cell_h;
        a->lists_capacity = grid->lists_capacity;
        memcpy(a->depth_split, p->depth_split, sizeof a->depth_split);
        const double ln = log2((double)p->near_plane), lf = log2((double)p->far_plane);
        const double scale = (double)p->grid_z / (lf - ln);
        a->log_scale = (float)scale;
        a->log_bias = (float)(-ln * scale);
    } else {
        return fail(c, HL_EINVAL, "unknown grid mode %u", grid->mode);
    }
    a->lists = grid->lists_d;
    a->counts = grid->counts_d;
    a->offsets = grid->offsets_d;
    a->cam = camera_d;
    a->width = W;
    a->height = H;
    a->rcp_w = 1.0f / (float)W;  // IEEE division: the correctly rounded reciprocal
    a->rcp_h = 1.0f / (float)H;
    a->albedo = {(const uint8_t*)gb->albedo.data_d, gb->albedo.pitch_bytes};
    a->normal = {(const uint8_t*)gb->normal.data_d, gb->normal.pitch_bytes};
    a->orm = {(const uint8_t*)gb->orm.data_d, gb->orm.pitch_bytes};
    a->emissive = {(const uint8_t*)gb->emissive.data_d, gb->emissive.pitch_bytes};
    a->depth = {(const uint8_t*)gb->depth.data_d, gb->depth.pitch_bytes};
    a->num_lights = num_lights;
    for (int k = 0; k < 3; ++k) {
#pragma clang fp contract(off)
        a->sky[k] = g->ambient_color[k] * g->ambient_intensity;
    }
    a->debug_mode = g->debug_mode;
    set_tile_shape(*a, 3);  // 8x8 pixels per wavefront
    return HL_OK;
}

}  // namespace hl

namespace {

// Shared set-up of the shading entry points: arguments, then the light records
// (built here, or with reuse those the context's last light cull built from
// the same lights).
// defer_prep: the records are built by the caller's own launch (into
// c->hot, for these lights) instead of launch_prep here.
hl_status prepare_shade(hl_context* c, const GpuCameraUniforms* camera_d, const DeferredGlobals* g,
                        const hl_gbuffer* gb, const GpuLight* lights_d, uint32_t num_lights,
                        const hl_light_grid* grid, ShadeArgs* a, int* set, bool need_debug0,
                        bool reuse = false, bool defer_prep = false) {
    if (!c) return HL_EINVAL;
    if (num_lights && !lights_d) return fail(c, HL_EINVAL, "lights are null");
    hl_status s = build_shade_args(c, camera_d, g, gb, num_lights, grid, a, set, need_debug0);
    if (s != HL_OK) return s;
    if ((s = ensure(c, c->words, 64)) != HL_OK) return s;
    HL_HIP(c, hipSetDevice(c->device));
    if (reuse) {
        if (!c->hot.p || c->hot_src != lights_d || c->hot_n != num_lights)
            return fail(c, HL_EINVAL, "HL_SHADE_LIGHTS_FROM_CULL: the last light cull on this context was "
                                      "not over these %u lights", num_lights);
        a->lights = (const HotLight*)c->hot.p;
        return HL_OK;
    }
    c->hot_src = nullptr;  // set again once the records are rebuilt
    if ((s = ensure(c, c->hot, sizeof(HotLight) * (size_t)(num_lights ? num_lights : 1))) != HL_OK) return s;
    a->lights = (const HotLight*)c->hot.p;
    if (!defer_prep)
        HL_HIP(c, launch_prep(c->stream, lights_d, nullptr, num_lights, camera_d, (HotLight*)c->hot.p,
                              nullptr, ctx_words(c) + 1));
    c->hot_src = lights_d;
    c->hot_n = num_lights;
    return HL_OK;
}

// Persistent shading launch on the context stream with the claim counters of
// the current parity; the other half is zeroed by the launch for the next one.
hl_status shade_launch(hl_context* c, int set, int mode, const ShadeArgs& a) {
    hl_status s = ensure(c, c->shadeq, 2 * 4 * (size_t)kShadeQueueWords, true);
    if (s != HL_OK) return s;
    uint32_t* q = (uint32_t*)c->shadeq.p;
    uint32_t* use = q + (size_t)c->shade_parity * kShadeQueueWords;
    uint32_t* next = q + (size_t)(c->shade_parity ^ 1u) * kShadeQueueWords;
    HL_HIP(c, launch_shade(c->stream, set, mode, a, use, next));
    c->shade_parity ^= 1u;
    return HL_OK;
}

}  // namespace

// ------------------------------------------- scratch of the no-alloc entry
// hl_deferred_forward_noalloc keeps its scratch -- two sets of claim counters
// and the shading records of up to HL_NOALLOC_MAX_LIGHTS lights -- in this
// file-scope device array rather than in context allocations.  Each device
// holds its own copy until the process ends.  Host code never uses the
// array's name as an address: it uses the device address hipGetSymbolAddress
// returns for the current device, looked up once per device (noalloc_setup,
// from hl_context_create) and kept in g_scratch_ptr and the context.
constexpr size_t kNoAllocLightsOff = 2 * 4 * (size_t)kShadeQueueWords;
constexpr size_t kNoAllocBytes = kNoAllocLightsOff + sizeof(HotLight) * (size_t)HL_NOALLOC_MAX_LIGHTS;
static_assert(kNoAllocLightsOff % 256 == 0, "light records start 256-B aligned");
static_assert(kNoAllocBytes == HL_NOALLOC_SCRATCH_BYTES, "HL_NOALLOC_SCRATCH_BYTES states the array's size");
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device ordinal; null until looked up

// The device of c is current.
hl_status noalloc_setup(hl_context* c) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)c->device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)c->device + 1, nullptr);
    if (!g_scratch_ptr[(size_t)c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                   "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[(size_t)c->device] = p;
    }
    c->noalloc = g_scratch_ptr[(size_t)c->device];
    return HL_OK;
}

}  // namespace

// ------------------------------------------------------------------ exports
extern "C" {

HL_API const char* hl_version(void) { return "helio_hip 0.1.0 (gfx950)"; }

HL_API const char* hl_status_string(hl_status s) {
    switch (s) {
        case HL_OK: return "ok";
        case HL_EINVAL: return "invalid argument";
        case HL_EHIP: return "HIP runtime error";
        case HL_ECAPACITY: return "capacity exceeded";
        case HL_EUNSUPPORTED: return "unsupported input";
        case HL_ENOMEM: return "out of device memory";
        default: return "unknown status";
    }
}

HL_API hl_status hl_context_create(int device, hl_context** out) {
    if (!out) return HL_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return HL_EHIP;
    if (device < 0 || device >= n) return HL_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return HL_EHIP;
    hl_context* c = new hl_context();
    c->device = device;
    if (ensure(c, c->words, 64) != HL_OK) {
        delete c;
        return HL_ENOMEM;
    }
    if (hipMemset(c->words.p, 0, c->words.bytes) != hipSuccess || noalloc_setup(c) != HL_OK) {
        (void)hipFree(c->words.p);
        delete c;
        return HL_EHIP;
    }
    *out = c;
    return HL_OK;
}

HL_API hl_status hl_context_destroy(hl_context* c) {
    if (!c) return HL_EINVAL;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    while (!c->externals.empty()) (void)release_external(c->externals.back());
    for (DevBuf* b : {&c->hot, &c->cull, &c->geom, &c->bits, &c->summary, &c->raw, &c->markq, &c->shadeq,
                      &c->words, &c->accum, &c->vogel, &c->rcirr, &c->tilebin})
        if (b->p) (void)hipFree(b->p);
    delete c;
    return HL_OK;
}

HL_API hl_status hl_set_stream(hl_context* c, void* stream) {
    if (!c) return HL_EINVAL;
    // work of the old stream finishes first: the context's scratch (and the
    // shading claim counters a launch leaves for the next one) is reused in
    // stream order
    HL_HIP(c, hipStreamSynchronize(c->stream));
    c->stream = (hipStream_t)stream;
    return HL_OK;
}

HL_API void* hl_get_stream(const hl_context* c) { return c ? (void*)c->stream : nullptr; }

HL_API hl_status hl_synchronize(hl_context* c) {
    if (!c) return HL_EINVAL;
    HL_HIP(c, hipStreamSynchronize(c->stream));
    return HL_OK;
}

HL_API const char* hl_last_error(const hl_context* c) {
    return c ? c->err.c_str() : "null context";
}

HL_API hl_status hl_alloc(hl_context* c, size_t bytes, void** out) {
    if (!c || !out) return HL_EINVAL;
    HL_HIP(c, hipSetDevice(c->device));
    hipError_t e = hipMalloc(out, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(c, HL_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return HL_OK;
}

HL_API hl_status hl_free(hl_context* c, void* p) {
    if (!c) return HL_EINVAL;
    if (!p) return HL_OK;
    HL_HIP(c, hipSetDevice(c->device));
    HL_HIP(c, hipStreamSynchronize(c->stream));
    HL_HIP(c, hipFree(p));
    return HL_OK;
}

HL_API hl_status hl_memcpy_htod(hl_context* c, void* dst, const void* src, size_t n) {
    if (!c || (n && (!dst || !src))) return HL_EINVAL;
    HL_HIP(c, hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream));
    HL_HIP(c, hipStreamSynchronize(c->stream));
    return HL_OK;
}

HL_API hl_status hl_memcpy_dtoh(hl_context* c, void* dst, const void* src, size_t n) {
    if (!c || (n && (!dst || !src))) return HL_EINVAL;
    HL_HIP(c, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->stream));
    HL_HIP(c, hipStreamSynchronize(c->stream));
    return HL_OK;
}

HL_API hl_status hl_memcpy_dtod(hl_context* c, void* dst, const void* src, size_t n) {
    if (!c || (n && (!dst || !src))) return HL_EINVAL;
    HL_HIP(c, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, c->stream));
    return HL_OK;
}

HL_API hl_status hl_memset_d32(hl_context* c, void* dst, uint32_t v, size_t count) {
    if (!c || (count && !dst)) return HL_EINVAL;
    HL_HIP(c, launch_fill_u32(c->stream, (uint32_t*)dst, v, count));
    return HL_OK;
}

HL_API hl_status hl_event_create(hl_context* c, hl_event** out) {
    if (!c || !out) return HL_EINVAL;
    hl_event* e = new hl_event();
    hipError_t r = hipEventCreate(&e->ev);
    if (r != hipSuccess) {
        delete e;
        return hip_fail(c, r, "hipEventCreate");
    }
    *out = e;
    return HL_OK;
}

HL_API hl_status hl_event_destroy(hl_event* e) {
    if (!e) return HL_EINVAL;
    (void)hipEventDestroy(e->ev);
    delete e;
    return HL_OK;
}

HL_API hl_status hl_event_record(hl_context* c, hl_event* e) {
    if (!c || !e) return HL_EINVAL;
    HL_HIP(c, hipEventRecord(e->ev, c->stream));
    return HL_OK;
}

HL_API hl_status hl_event_elapsed_ms(hl_event* a, hl_event* b, float* ms) {
    if (!a || !b || !ms) return HL_EINVAL;
    if (hipEventSynchronize(b->ev) != hipSuccess) return HL_EHIP;
    if (hipEventElapsedTime(ms, a->ev, b->ev) != hipSuccess) return HL_EHIP;
    return HL_OK;
}

HL_API hl_status hl_cluster_params_init(const GpuCameraUniforms* cam, uint32_t width,
                                        uint32_t height, uint32_t gx, uint32_t gy, uint32_t gz,
                                        uint32_t cap, hl_cluster_params* out) {
    if (!cam || !out) return HL_EINVAL;
    if (width == 0 || height == 0 || gx == 0 || gy == 0 || gz == 0 || gz > HL_MAX_SLICES)
        return HL_EINVAL;
    // standard-Z perspective: proj[2][3] = -1, proj[3][3] = 0, no off-centre terms
    if (cam->proj[11] != -1.0f || cam->proj[15] != 0.0f || cam->proj[8] != 0.0f ||
        cam->proj[9] != 0.0f)
        return HL_EUNSUPPORTED;
    const double p22 = cam->proj[10], p32 = cam->proj[14];
    const double near_ = cam->position_near[3], far_ = cam->forward_far[3];
    if (!(near_ > 0.0) || !(far_ > near_)) return HL_EINVAL;
    memset(out, 0, sizeof *out);
    out->grid_x = gx;
    out->grid_y = gy;
    out->grid_z = gz;
    out->cell_w = (width + gx - 1) / gx;
    out->cell_h = (height + gy - 1) / gy;
    out->width = width;
    out->height = height;
    out->cap = cap;
    out->near_plane = (float)near_;
    out->far_plane = (float)far_;
    // NDC depth of view distance z: (p22 * (-z) + p32) / z; view distance of depth d: p32 / (d + p22)
    for (uint32_t k = 0; k <= gz; ++k) {
        const double z = near_ * pow(far_ / near_, (double)k / gz);
        double d = (p22 * (-z) + p32) / z;
        if (k == 0) d = 0.0;
        if (k == gz) d = 1.0;
        out->depth_split[k] = (float)d;
    }
    // View-distance range of each slice, padded for the FP32 reconstruction
    // error of a world position from its depth: w = m23 * depth + m33 with
    // |m33| = 1/near cancels to 1/distance, so w carries an absolute error of
    // ~1.2e-7/near, a relative error of ~1.2e-7 * distance / near (x2.5 margin
    // below, plus 1e-5 for the other rounding steps).  The radius padding
    // bounds the same error as a world-space distance (twice the relative
    // error times the point's distance from the origin, plus 1 mm).
    const double cam_dist = sqrt((double)cam->position_near[0] * cam->position_near[0] +
                                 (double)cam->position_near[1] * cam->position_near[1] +
                                 (double)cam->position_near[2] * cam->position_near[2]);
    double maxpad = 0.0;
    for (uint32_t k = 0; k < gz; ++k) {
        const double d0 = out->depth_split[k], d1 = out->depth_split[k + 1];
        double z0 = k == 0 ? near_ : p32 / (d0 + p22);
        double z1 = k + 1 == gz ? far_ : p32 / (d1 + p22);
        const double rel = 1e-5 + 3e-7 * z1 / near_;
        z0 *= (1.0 - rel);
        z1 *= (1.0 + rel);
        out->slice_near[k] = (float)z0;
        out->slice_far[k] = (float)z1;
        const double pad = 2.0 * rel * (z1 + cam_dist) + 1e-3;
        out->slice_r_pad[k] = (float)pad;
        if (pad > maxpad) maxpad = pad;
    }
    // Non-decreasing tables (the cull searches them): a slice's near limit is
    // lowered to the next one's, its far limit and padding raised to the
    // previous one's where FP32 depths round to equal distances -- wider, so
    // still conservative.
    for (uint32_t k = gz - 1; k-- > 0;)
        out->slice_near[k] = std::min(out->slice_near[k], out->slice_near[k + 1]);
    for (uint32_t k = 1; k < gz; ++k) {
        out->slice_far[k] = std::max(out->slice_far[k], out->slice_far[k - 1]);
        out->slice_r_pad[k] = std::max(out->slice_r_pad[k], out->slice_r_pad[k - 1]);
    }
    out->r_pad_column = (float)maxpad;
    return HL_OK;
}

HL_API hl_status hl_light_cull_tiles(hl_context* c, const GpuCameraUniforms* camera_d,
                                     const LightCullParams* params, const GpuLight* lights_d,
                                     uint32_t* lists_d, uint32_t* counts_d) {
    if (!c) return HL_EINVAL;
    if (!params || !lists_d || !counts_d) return fail(c, HL_EINVAL, "null params or outputs");
    const uint32_t tiles = params->num_tiles_x * params->num_tiles_y;
    HL_HIP(c, hipSetDevice(c->device));
    if (params->num_lights == 0) {  // LightCullPass::execute, lib.rs:285-292
        HL_HIP(c, launch_fill_u32(c->stream, lists_d, 0, (size_t)tiles * HL_MAX_LIGHTS_PER_TILE));
        HL_HIP(c, launch_fill_u32(c->stream, counts_d, 0, tiles));
        return HL_OK;
    }
    if (!camera_d || !lights_d) return fail(c, HL_EINVAL, "null camera or lights");
    if (params->screen_width == 0 || params->screen_height == 0)
        return fail(c, HL_EINVAL, "zero screen size");
    c->hot_src = nullptr;  // set again once the records are rebuilt
    hl_status s = ensure(c, c->cull, sizeof(CullLight) * (size_t)params->num_lights);
    if (s != HL_OK) return s;
    if ((s = ensure(c, c->hot, sizeof(HotLight) * (size_t)params->num_lights)) != HL_OK) return s;
    HL_HIP(c, launch_prep(c->stream, lights_d, nullptr, params->num_lights, camera_d,
                          (HotLight*)c->hot.p, (CullLight*)c->cull.p, ctx_words(c) + 1));
    c->hot_src = lights_d;
    c->hot_n = params->num_lights;
    const LightCullParams& p = *params;
    const uint64_t n = p.num_lights, nsup = tile_bin_supers(p);
    if (p.num_tiles_x == (p.screen_width + HL_TILE_SIZE - 1) / HL_TILE_SIZE &&
        p.num_tiles_y == (p.screen_height + HL_TILE_SIZE - 1) / HL_TILE_SIZE && nsup * n <= 0xFFFFFFFFull) {
        // the screen's tile grid: candidates binned per super-tile
        if ((s = ensure(c, c->bits, 4 * cluster_bits_words((uint32_t)nsup, (uint32_t)n), true)) != HL_OK) return s;
        if ((s = ensure(c, c->summary, 4 * cluster_summary_words((uint32_t)nsup, (uint32_t)n), true)) != HL_OK)
            return s;
        if ((s = ensure(c, c->raw, cluster_raw_bytes((uint32_t)nsup))) != HL_OK) return s;
        const uint64_t tiles_n = (uint64_t)p.num_tiles_x * p.num_tiles_y;
        if ((s = ensure(c, c->tilebin, 4 * (12 * tiles_n + 2 * nsup + nsup * n))) != HL_OK) return s;
        float* planes = (float*)c->tilebin.p;
        uint32_t* tb = (uint32_t*)(planes + 12 * tiles_n);
        const TileBinScratch w{(uint32_t*)c->bits.p, (uint32_t*)c->summary.p, (uint32_t*)c->raw.p, tb, tb + nsup,
                               tb + 2 * nsup, (uint32_t)(nsup * n), ctx_words(c) + 4, planes};
        HL_HIP(c, launch_cull_tiles_binned(c->stream, camera_d, p, (const CullLight*)c->cull.p, lists_d, counts_d,
                                           w));
    } else {
        HL_HIP(c, launch_cull_tiles(c->stream, camera_d, p, (const CullLight*)c->cull.p, lists_d, counts_d));
    }
    return HL_OK;
}

HL_API hl_status hl_light_cull_clusters(hl_context* c, const GpuCameraUniforms* camera_d,
                                        const hl_cluster_params* p, const GpuLight* lights_d,
                                        uint32_t num_lights, uint32_t* offsets_d,
                                        uint32_t* counts_d, uint32_t* pool_d, uint32_t pool_cap,
                                        uint32_t* overflow_d) {
    if (!c) return HL_EINVAL;
    hl_status s = check_cluster_params(c, p);
    if (s != HL_OK) return s;
    if (!camera_d) return fail(c, HL_EINVAL, "camera is null");
    if (num_lights && !lights_d) return fail(c, HL_EINVAL, "lights are null");
    if (!offsets_d || !counts_d || (!pool_d && pool_cap)) return fail(c, HL_EINVAL, "null outputs");
    const uint32_t ncol = p->grid_x * p->grid_y;
    const size_t ncell = (size_t)ncol * p->grid_z;
    const uint32_t n = num_lights;
    c->hot_src = nullptr;  // set again once the records are rebuilt
    if ((s = ensure(c, c->cull, sizeof(CullLight) * (size_t)(n ? n : 1))) != HL_OK) return s;
    if ((s = ensure(c, c->hot, sizeof(HotLight) * (size_t)(n ? n : 1))) != HL_OK) return s;
    if ((s = ensure(c, c->geom, cluster_geom_bytes(ncol, (uint32_t)ncell))) != HL_OK) return s;
    if ((s = ensure(c, c->bits, 4 * cluster_bits_words(ncell, n), true)) != HL_OK) return s;
    if ((s = ensure(c, c->summary, 4 * cluster_summary_words(ncell, n), true)) != HL_OK) return s;
    if ((s = ensure(c, c->raw, cluster_raw_bytes((uint32_t)ncell))) != HL_OK) return s;
    if ((s = ensure(c, c->markq, mark_queue_bytes(n))) != HL_OK) return s;
    HL_HIP(c, hipSetDevice(c->device));
    // the overflow word(s) are reset by the first cull kernel and OR-ed by the
    // list kernel: the caller's word when given, and always the context's copy
    uint32_t* flags = overflow_d ? overflow_d : ctx_words(c);
    uint32_t* flags2 = overflow_d ? ctx_words(c) : nullptr;
    ClusterArgs a = make_cluster_args(p, n);
    ClusterScratch w{c->geom.p, (uint32_t*)c->bits.p, (uint32_t*)c->summary.p, (uint32_t*)c->raw.p,
                     mark_queue_at(c->markq.p, n)};
    HL_HIP(c, launch_cluster_cull(c->stream, a, camera_d, lights_d, (HotLight*)c->hot.p, (CullLight*)c->cull.p,
                                  ctx_words(c) + 1, w, offsets_d, counts_d, pool_d, pool_cap, flags, flags2));
    c->hot_src = lights_d;
    c->hot_n = n;
    return HL_OK;
}

HL_API hl_status hl_query_overflow(hl_context* c, uint32_t* flags) {
    if (!c || !flags) return HL_EINVAL;
    HL_HIP(c, hipMemcpyAsync(flags, ctx_words(c), 4, hipMemcpyDeviceToHost, c->stream));
    HL_HIP(c, hipStreamSynchronize(c->stream));
    return HL_OK;
}

HL_API hl_status hl_gbuffer_pack(hl_context* c, const hl_surface* sf, const hl_gbuffer* g) {
    if (!c) return HL_EINVAL;
    if (!sf || !g) return fail(c, HL_EINVAL, "surface or gbuffer is null");
    if (sf->width != g->width || sf->height != g->height)
        return fail(c, HL_EINVAL, "surface %ux%u differs from gbuffer %ux%u", sf->width, sf->height, g->width,
                    g->height);
    if (g->width > 32768 || g->height > 32768)
        return fail(c, HL_EINVAL, "gbuffer %ux%u exceeds 32768", g->width, g->height);
    if (sf->pitch_bytes % 16 || (uint64_t)sf->pitch_bytes < 16ull * sf->width)
        return fail(c, HL_EINVAL, "surface pitch %u invalid for width %u (RGBA32F planes)", sf->pitch_bytes,
                    sf->width);
    const float* planes[5] = {sf->albedo_d, sf->normal_d, sf->orm_d, sf->emissive_d, sf->specular_f0_d};
    for (int i = 0; i < 5; ++i) {
        if (!planes[i] && i < 4) return fail(c, HL_EINVAL, "surface plane %d is null", i);
        if (((uintptr_t)planes[i]) % 16) return fail(c, HL_EINVAL, "surface plane %d not 16-B aligned", i);
    }
    hl_status s;
    if ((s = check_tex(c, g->albedo, g->width, g->height, "albedo")) != HL_OK) return s;
    if ((s = check_tex(c, g->normal, g->width, g->height, "normal")) != HL_OK) return s;
    if ((s = check_tex(c, g->orm, g->width, g->height, "orm")) != HL_OK) return s;
    if ((s = check_tex(c, g->emissive, g->width, g->height, "emissive")) != HL_OK) return s;
    int set = 0;
    if ((s = color_set(c, g, &set)) != HL_OK) return s;
    if (g->lightmap_uv.data_d) {
        if (g->lightmap_uv.format != HL_FMT_RG16_FLOAT)
            return fail(c, HL_EINVAL, "gbuffer.lightmap_uv must be RG16_FLOAT");
        if ((s = check_tex(c, g->lightmap_uv, g->width, g->height, "lightmap_uv")) != HL_OK) return s;
    }
    HL_HIP(c, hipSetDevice(c->device));
    HL_HIP(c, launch_gbuffer_pack(c->stream, set, *sf, *g));
    return HL_OK;
}

}  // extern "C"

namespace {

// hl_deferred_forward with optional shadows (sh) and radiance-cascade GI (rc).
hl_status run_forward(hl_context* c, const GpuCameraUniforms* camera_d, const DeferredGlobals* g,
                      const hl_gbuffer* gb, const GpuLight* lights_d, uint32_t num_lights, const hl_light_grid* grid,
                      const hl_shadow_maps* sh, const hl_texture* rc, float* out_d, uint32_t out_pitch_bytes,
                      uint32_t flags) {
    if (!c) return HL_EINVAL;
    if (!out_d) return fail(c, HL_EINVAL, "output is null");
    ShadeArgs a;
    int set;
    if (flags & ~(uint32_t)(HL_SHADE_CLEAR_SKY | HL_SHADE_LIGHTS_FROM_CULL | HL_SHADE_OUT_RGBA16F))
        return fail(c, HL_EINVAL, "unknown shading flags 0x%x", flags);
    if (sh) {
        if (sh->atlas_size < 2 || sh->atlas_size > 16384)
            return fail(c, HL_EINVAL, "shadow atlas_size %u out of range (2..16384)", sh->atlas_size);
        if (sh->num_layers > 65536) return fail(c, HL_EINVAL, "%u shadow layers (at most 65536)", sh->num_layers);
        if (sh->num_layers && (!sh->atlas_d || !sh->static_atlas_d || !sh->matrices_d))
            return fail(c, HL_EINVAL, "shadow atlases or matrices are null");
        if (((uintptr_t)sh->atlas_d | (uintptr_t)sh->static_atlas_d | (uintptr_t)sh->matrices_d) % 4)
            return fail(c, HL_EINVAL, "shadow atlases and matrices must be 4-byte aligned");
        const ShadowConfig& k = sh->config;
        if (k.pcf_sample_count > 256 || k.pcss_blocker_samples > 256 || k.pcss_filter_samples > 256)
            return fail(c, HL_EINVAL, "shadow sample counts above 256");
    }
    const bool use_rc = rc && rc->data_d && g && g->has_rc_gi != 0u;
    if (use_rc) {
        if (rc->format != HL_FMT_RGBA16_FLOAT && rc->format != HL_FMT_RGBA32_FLOAT)
            return fail(c, HL_EINVAL, "rc_cascade0 must be RGBA16_FLOAT or RGBA32_FLOAT");
        // 16-byte rows and start: the kernel reads two RGBA16F texels per load
        const uint32_t tb = rc->format == HL_FMT_RGBA16_FLOAT ? 8u : 16u;
        if (rc->pitch_bytes < 64u * tb || rc->pitch_bytes % 16u || ((uintptr_t)rc->data_d) % 16u)
            return fail(c, HL_EINVAL, "rc_cascade0: pitch %u invalid for 64 texels or not 16-byte aligned",
                        rc->pitch_bytes);
    }
    hl_status s = prepare_shade(c, camera_d, g, gb, lights_d, num_lights, grid, &a, &set, false,
                                (flags & HL_SHADE_LIGHTS_FROM_CULL) != 0);
    if (s != HL_OK) return s;
    const bool f16 = (flags & HL_SHADE_OUT_RGBA16F) != 0u;
    const uint32_t px = f16 ? 8u : 16u;
    if ((uint64_t)out_pitch_bytes < (uint64_t)px * gb->width || out_pitch_bytes % px != 0 ||
        ((uintptr_t)out_d) % px)
        return fail(c, HL_EINVAL, "output pitch %u invalid for width %u (%s, %u-B aligned)", out_pitch_bytes,
                    gb->width, f16 ? "RGBA16F" : "RGBA32F", px);
    a.out = out_d;
    a.out_pitch = out_pitch_bytes;
    a.out_f16 = f16 ? 1u : 0u;
    a.flags = flags;
    int mode = kModeForward;
    if (use_rc) {
        a.rc = Tex{(const uint8_t*)rc->data_d, rc->pitch_bytes};
        a.rc_fmt = rc->format;
        for (int k = 0; k < 3; ++k) {
            a.rc_min[k] = g->rc_world_min[k];
            a.rc_size[k] = g->rc_world_max[k] - g->rc_world_min[k];
        }
        mode = kModeExtended;  // shadows (if any) and RC
        if ((s = ensure(c, c->rcirr, 16 * (size_t)gb->width * gb->height)) != HL_OK) return s;
        a.rc_irr = (const float4*)c->rcirr.p;
    }
    if (sh) {
        ShadowView& v = a.shadow;
        v.atlas = sh->atlas_d;
        v.static_atlas = sh->static_atlas_d;
        v.matrices = sh->matrices_d;
        v.layers = sh->num_layers;
        v.size = sh->atlas_size;
        v.pcf_count = sh->config.pcf_sample_count;
        v.enable_pcss = sh->config.enable_pcss;
        v.blocker_samples = sh->config.pcss_blocker_samples;
        v.filter_samples = sh->config.pcss_filter_samples;
        for (int k = 0; k < 4; ++k) {
            v.filter_radius[k] = sh->config.cascades[k].filter_radius;
            v.light_size[k] = sh->config.cascades[k].pcss_light_size;
            v.csm[k] = g->csm_splits[k];
        }
        if ((s = ensure(c, c->vogel, sizeof(float) * kVogelTableFloats)) != HL_OK) return s;
        if (!c->vogel_ready) {
            HL_HIP(c, launch_vogel_table(c->stream, (float*)c->vogel.p));
            c->vogel_ready = true;
        }
        v.vogel = (const float*)c->vogel.p;
        if (mode != kModeExtended) mode = kModeShadowed;
    }
    if (mode == kModeExtended) HL_HIP(c, launch_rc_irradiance(c->stream, set, a));
    return shade_launch(c, set, mode, a);
}

}  // namespace

extern "C" {

HL_API hl_status hl_deferred_forward(hl_context* c, const GpuCameraUniforms* camera_d,
                                     const DeferredGlobals* g, const hl_gbuffer* gb,
                                     const GpuLight* lights_d, uint32_t num_lights,
                                     const hl_light_grid* grid, float* out_d,
                                     uint32_t out_pitch_bytes, uint32_t flags) {
    return run_forward(c, camera_d, g, gb, lights_d, num_lights, grid, nullptr, nullptr, out_d, out_pitch_bytes,
                       flags);
}

HL_API hl_status hl_deferred_forward_noalloc(hl_context* c, const GpuCameraUniforms* camera_d,
                                             const DeferredGlobals* g, const hl_gbuffer* gb,
                                             const GpuLight* lights_d, uint32_t num_lights,
                                             const hl_light_grid* grid, float* out_d,
                                             uint32_t out_pitch_bytes, uint32_t flags) {
    if (!c) return HL_EINVAL;
    if (!c->noalloc) return fail(c, HL_EHIP, "the no-alloc scratch of device %d is not set up", c->device);
    if (!out_d) return fail(c, HL_EINVAL, "output is null");
    if (flags & ~(uint32_t)HL_SHADE_CLEAR_SKY)
        return fail(c, HL_EINVAL, "hl_deferred_forward_noalloc: flags 0x%x (HL_SHADE_CLEAR_SKY only)", flags);
    if (num_lights > HL_NOALLOC_MAX_LIGHTS)
        return fail(c, HL_EINVAL, "hl_deferred_forward_noalloc: %u lights, at most %u (HL_NOALLOC_MAX_LIGHTS)",
                    num_lights, HL_NOALLOC_MAX_LIGHTS);
    if (num_lights && !lights_d) return fail(c, HL_EINVAL, "lights are null");
    ShadeArgs a;
    int set;
    hl_status s = build_shade_args(c, camera_d, g, gb, num_lights, grid, &a, &set, false);
    if (s != HL_OK) return s;
    if (out_pitch_bytes < 16u * gb->width || out_pitch_bytes % 16 != 0 || ((uintptr_t)out_d) % 16)
        return fail(c, HL_EINVAL, "output pitch %u invalid for width %u (RGBA32F, 16-B aligned)",
                    out_pitch_bytes, gb->width);
    a.out = out_d;
    a.out_pitch = out_pitch_bytes;
    a.flags = flags;
    HL_HIP(c, hipSetDevice(c->device));
    uint8_t* base = (uint8_t*)c->noalloc;  // this device's address of `scratch`
    uint32_t* queue = (uint32_t*)base;
    HotLight* hot = (HotLight*)(base + kNoAllocLightsOff);
    a.lights = hot;
    HL_HIP(c, launch_prep(c->stream, lights_d, nullptr, num_lights, camera_d, hot, nullptr, nullptr));
    // the array keeps the last call's counters: clear this launch's set (a
    // fill kernel, so that the call can be captured into a graph)
    HL_HIP(c, launch_fill_u32(c->stream, queue, 0, kShadeQueueWords));
    HL_HIP(c, launch_shade(c->stream, set, kModeForward, a, queue, queue + kShadeQueueWords));
    return HL_OK;
}

HL_API hl_status hl_deferred_forward_shadowed(hl_context* c, const GpuCameraUniforms* camera_d,
                                              const DeferredGlobals* g, const hl_gbuffer* gb,
                                              const GpuLight* lights_d, uint32_t num_lights,
                                              const hl_light_grid* grid, const hl_shadow_maps* shadows,
                                              float* out_d, uint32_t out_pitch_bytes, uint32_t flags) {
    return run_forward(c, camera_d, g, gb, lights_d, num_lights, grid, shadows, nullptr, out_d, out_pitch_bytes,
                       flags);
}

HL_API hl_status hl_deferred_forward_ex(hl_context* c, const GpuCameraUniforms* camera_d, const DeferredGlobals* g,
                                        const hl_gbuffer* gb, const GpuLight* lights_d, uint32_t num_lights,
                                        const hl_light_grid* grid, const hl_shadow_maps* shadows,
                                        const hl_texture* rc_cascade0, float* out_d, uint32_t out_pitch_bytes,
                                        uint32_t flags) {
    return run_forward(c, camera_d, g, gb, lights_d, num_lights, grid, shadows, rc_cascade0, out_d,
                       out_pitch_bytes, flags);
}

HL_API hl_status hl_shade_walk_stats(hl_context* c, const GpuCameraUniforms* camera_d, const DeferredGlobals* g,
                                     const hl_gbuffer* gb, const GpuLight* lights_d, uint32_t num_lights,
                                     const hl_light_grid* grid, uint64_t* pairs, uint64_t* wave_light_visits) {
    if (!c) return HL_EINVAL;
    if (!pairs || !wave_light_visits) return fail(c, HL_EINVAL, "pairs or wave_light_visits is null");
    DeferredGlobals g0;
    if (!g) return fail(c, HL_EINVAL, "globals are null");
    g0 = *g;
    g0.debug_mode = 0;
    ShadeArgs a;
    int set;
    hl_status s = prepare_shade(c, camera_d, &g0, gb, lights_d, num_lights, grid, &a, &set, false);
    if (s != HL_OK) return s;
    // words 2-3: pairs, 6-7: visits (4-5, the binned tile cull's overflow
    // word, stay zero)
    unsigned long long* ctr = (unsigned long long*)(ctx_words(c) + 2);
    HL_HIP(c, launch_fill_u32(c->stream, (uint32_t*)ctr, 0, 6));
    a.pair_counter = ctr;
    if ((s = shade_launch(c, set, kModeCount, a)) != HL_OK) return s;
    unsigned long long v[3] = {0, 0, 0};
    HL_HIP(c, hipMemcpyAsync(v, ctr, sizeof v, hipMemcpyDeviceToHost, c->stream));
    HL_HIP(c, hipStreamSynchronize(c->stream));
    *pairs = v[0];
    *wave_light_visits = v[2];
    return HL_OK;
}

HL_API hl_status hl_count_pixel_light_pairs(hl_context* c, const GpuCameraUniforms* camera_d,
                                            const DeferredGlobals* g, const hl_gbuffer* gb,
                                            const GpuLight* lights_d, uint32_t num_lights,
                                            const hl_light_grid* grid, uint64_t* pairs) {
    if (!c) return HL_EINVAL;
    if (!pairs) return fail(c, HL_EINVAL, "pairs is null");
    uint64_t visits = 0;
    return hl_shade_walk_stats(c, camera_d, g, gb, lights_d, num_lights, grid, pairs, &visits);
}

}  // extern "C"

namespace {

hl_status check_rgba32(hl_context* c, const void* p, uint32_t pitch, uint32_t w, const char* what) {
    if (!p) return fail(c, HL_EINVAL, "%s is null", what);
    if (pitch < 16u * w || pitch % 16 != 0 || ((uintptr_t)p) % 16)
        return fail(c, HL_EINVAL, "%s: pitch %u invalid for width %u (RGBA32F, 16-B aligned)", what,
                    pitch, w);
    return HL_OK;
}

hl_status run_adjoint(hl_context* c, const GpuCameraUniforms* camera_d, const DeferredGlobals* g,
                      const hl_gbuffer* gb, const GpuLight* lights_d, uint32_t n,
                      const hl_light_grid* grid, const float* in_d, uint32_t in_pitch,
                      float* out_d, uint32_t out_pitch, double* loss_d, float* d_pixel_d,
                      float* d_pos_d, float* d_int_d, bool fused) {
    if (!c) return HL_EINVAL;
    ShadeArgs a;
    int set;
    // the records are built below, with the accumulators cleared, in one launch
    hl_status s = prepare_shade(c, camera_d, g, gb, lights_d, n, grid, &a, &set, true, false, true);
    if (s != HL_OK) return s;
    c->hot_src = nullptr;  // set again once the launch below is queued
    if ((s = check_rgba32(c, in_d, in_pitch, gb->width, fused ? "target" : "dL_dout")) != HL_OK) return s;
    if (fused) {
        if ((s = check_rgba32(c, out_d, out_pitch, gb->width, "output")) != HL_OK) return s;
        if (!loss_d) return fail(c, HL_EINVAL, "loss is null");
    }
    if (!d_pixel_d) return fail(c, HL_EINVAL, "d_pixel is null");
    if (n && (!d_pos_d || !d_int_d)) return fail(c, HL_EINVAL, "light gradient outputs are null");
    if ((s = ensure(c, c->accum, sizeof(double) * grad_accum_doubles(n))) != HL_OK) return s;
    double* acc = (double*)c->accum.p;
    HL_HIP(c, launch_prep_zero(c->stream, lights_d, n, camera_d, (HotLight*)c->hot.p, ctx_words(c) + 1, acc,
                               grad_accum_doubles(n), fused ? loss_d : nullptr));
    c->hot_src = lights_d;
    c->hot_n = n;
    a.dout = in_d;
    a.dout_pitch = in_pitch;
    a.out = out_d;
    a.out_pitch = out_pitch;
    a.loss = loss_d;
    a.d_pixel = d_pixel_d;
    a.d_light = acc;
    if ((s = shade_launch(c, set, fused ? kModeFusedL2 : kModeBackward, a)) != HL_OK) return s;
    HL_HIP(c, launch_light_grads_to_f32(c->stream, acc, n, nullptr, d_pos_d, d_int_d));
    return HL_OK;
}

}  // namespace

extern "C" {

HL_API hl_status hl_deferred_backward(hl_context* c, const GpuCameraUniforms* camera_d,
                                      const DeferredGlobals* g, const hl_gbuffer* gb,
                                      const GpuLight* lights_d, uint32_t num_lights,
                                      const hl_light_grid* grid, const float* dL_dout_d,
                                      uint32_t dout_pitch_bytes, float* d_pixel_d,
                                      float* d_light_pos_d, float* d_light_intensity_d) {
    return run_adjoint(c, camera_d, g, gb, lights_d, num_lights, grid, dL_dout_d, dout_pitch_bytes,
                       nullptr, 0, nullptr, d_pixel_d, d_light_pos_d, d_light_intensity_d, false);
}

HL_API hl_status hl_deferred_forward_backward_l2(
    hl_context* c, const GpuCameraUniforms* camera_d, const DeferredGlobals* g,
    const hl_gbuffer* gb, const GpuLight* lights_d, uint32_t num_lights, const hl_light_grid* grid,
    const float* target_d, uint32_t target_pitch_bytes, float* out_d, uint32_t out_pitch_bytes,
    double* loss_d, float* d_pixel_d, float* d_light_pos_d, float* d_light_intensity_d) {
    return run_adjoint(c, camera_d, g, gb, lights_d, num_lights, grid, target_d, target_pitch_bytes,
                       out_d, out_pitch_bytes, loss_d, d_pixel_d, d_light_pos_d, d_light_intensity_d,
                       true);
}

}  // extern "C"

// ---------------------------------------------------------------- forward-lit
namespace {

// Arguments of a forward-lit launch, and the light records it reads (built on
// the context stream from the lights and, when given, their transforms).
hl_status prepare_forward(hl_context* c, const GpuCameraUniforms* camera_d, const ForwardLitGlobals* g,
                          const hl_fragment* frags_d, uint32_t n, const GpuLight* lights_d,
                          uint32_t num_lights, const uint32_t* entity_d, const hl_transform* transforms_d,
                          uint32_t num_transforms, const hl_light_grid* grid, ForwardArgs* a) {
    if (!c) return HL_EINVAL;
    if (!camera_d) return fail(c, HL_EINVAL, "camera is null");
    if (!g) return fail(c, HL_EINVAL, "globals are null");
    if (!grid) return fail(c, HL_EINVAL, "light grid is null");
    if (n && !frags_d) return fail(c, HL_EINVAL, "fragments are null");
    if (((uintptr_t)frags_d) % 16) return fail(c, HL_EINVAL, "fragments must be 16-byte aligned");
    if (num_lights && !lights_d) return fail(c, HL_EINVAL, "lights are null");
    if (!entity_d != !transforms_d)
        return fail(c, HL_EINVAL, "light_entity_indices and transforms must both be given or both be null");
    memset(a, 0, sizeof *a);
    hl_status s;
    if (grid->mode == HL_GRID_TILES) {
        if (!grid->lists_d || !grid->counts_d) return fail(c, HL_EINVAL, "tile lists/counts are null");
        if (g->num_tiles_x != grid->num_tiles_x)
            return fail(c, HL_EINVAL, "globals.num_tiles_x (%u) != grid.num_tiles_x (%u)", g->num_tiles_x,
                        grid->num_tiles_x);
        a->grid_mode = HL_GRID_TILES;
        a->grid_x = grid->num_tiles_x;
        a->grid_y = grid->num_tiles_y;
        a->grid_z = 1;
        a->cell_w = a->cell_h = HL_TILE_SIZE;
        a->tile_stride_x = g->num_tiles_x;
        uint64_t cap = (uint64_t)grid->num_tiles_x * grid->num_tiles_y * HL_MAX_LIGHTS_PER_TILE;
        if (grid->lists_capacity && grid->lists_capacity < cap) cap = grid->lists_capacity;
        a->lists_capacity = cap > 0xffffffffull ? 0xffffffffu : (uint32_t)cap;
    } else if (grid->mode == HL_GRID_CLUSTERS) {
        const hl_cluster_params* p = grid->clusters;
        if ((s = check_cluster_params(c, p)) != HL_OK) return s;
        if (!grid->lists_d || !grid->counts_d || !grid->offsets_d)
            return fail(c, HL_EINVAL, "cluster offsets/counts/pool are null");
        a->grid_mode = HL_GRID_CLUSTERS;
        a->grid_x = p->grid_x;
        a->grid_y = p->grid_y;
        a->grid_z = p->grid_z;
        a->cell_w = p->cell_w;
        a->cell_h = p->cell_h;
        a->lists_capacity = grid->lists_capacity;
        memcpy(a->depth_split, p->depth_split, sizeof a->depth_split);
    } else {
        return fail(c, HL_EINVAL, "unknown grid mode %u", grid->mode);
    }
    a->lists = grid->lists_d;
    a->counts = grid->counts_d;
    a->offsets = grid->offsets_d;
    a->cam = camera_d;
    a->frags = frags_d;
    a->num_frags = n;
    a->num_lights = num_lights;
    for (int k = 0; k < 3; ++k) {
#pragma clang fp contract(off)
        a->amb[k] = g->ambient_color[k] * g->ambient_intensity;
    }
    if ((s = ensure(c, c->words, 64)) != HL_OK) return s;
    HL_HIP(c, hipSetDevice(c->device));
    c->hot_src = nullptr;  // the records no longer mirror a light cull's
    if ((s = ensure(c, c->hot, sizeof(HotLight) * (size_t)(num_lights ? num_lights : 1))) != HL_OK) return s;
    a->lights = (const HotLight*)c->hot.p;
    HL_HIP(c, launch_forward_prep(c->stream, lights_d, num_lights, entity_d, transforms_d, num_transforms,
                                  (HotLight*)c->hot.p));
    return HL_OK;
}

}  // namespace

extern "C" {

HL_API hl_status hl_forward_lit(hl_context* c, const GpuCameraUniforms* camera_d, const ForwardLitGlobals* g,
                                const hl_fragment* fragments_d, uint32_t num_fragments, const GpuLight* lights_d,
                                uint32_t num_lights, const uint32_t* light_entity_indices_d,
                                const hl_transform* transforms_d, uint32_t num_transforms,
                                const hl_light_grid* grid, float* out_d, uint32_t flags) {
    if (!c) return HL_EINVAL;
    if (flags) return fail(c, HL_EINVAL, "unknown forward-lit flags 0x%x", flags);
    if (num_fragments && !out_d) return fail(c, HL_EINVAL, "output is null");
    if (((uintptr_t)out_d) % 16) return fail(c, HL_EINVAL, "output must be 16-byte aligned");
    ForwardArgs a;
    hl_status s = prepare_forward(c, camera_d, g, fragments_d, num_fragments, lights_d, num_lights,
                                  light_entity_indices_d, transforms_d, num_transforms, grid, &a);
    if (s != HL_OK) return s;
    a.out = out_d;
    HL_HIP(c, launch_forward_lit(c->stream, a));
    return HL_OK;
}

HL_API hl_status hl_forward_lit_count_pairs(hl_context* c, const GpuCameraUniforms* camera_d,
                                            const ForwardLitGlobals* g, const hl_fragment* fragments_d,
                                            uint32_t num_fragments, const GpuLight* lights_d, uint32_t num_lights,
                                            const uint32_t* light_entity_indices_d,
                                            const hl_transform* transforms_d, uint32_t num_transforms,
                                            const hl_light_grid* grid, uint64_t* pairs) {
    if (!c) return HL_EINVAL;
    if (!pairs) return fail(c, HL_EINVAL, "pairs is null");
    ForwardArgs a;
    hl_status s = prepare_forward(c, camera_d, g, fragments_d, num_fragments, lights_d, num_lights,
                                  light_entity_indices_d, transforms_d, num_transforms, grid, &a);
    if (s != HL_OK) return s;
    unsigned long long* ctr = (unsigned long long*)(ctx_words(c) + 2);
    HL_HIP(c, launch_fill_u32(c->stream, (uint32_t*)ctr, 0, 2));
    a.pair_counter = ctr;
    HL_HIP(c, launch_forward_lit(c->stream, a));
    unsigned long long v = 0;
    HL_HIP(c, hipMemcpyAsync(&v, ctr, 8, hipMemcpyDeviceToHost, c->stream));
    HL_HIP(c, hipStreamSynchronize(c->stream));
    *pairs = v;
    return HL_OK;
}

}  // extern "C"
