// hl_internal.hpp -- declarations shared by the kernel and API translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "hl_device.hpp"

namespace hl {

// bits of the context's feature flag word (set by the prep kernel)
constexpr uint32_t kFlagShadowIgnored = 1u;   // a light < 42 has a shadow_index
constexpr uint32_t kFlagTexturedIgnored = 2u; // a light has an IES profile or gobo

struct ClusterArgs {
    uint32_t grid_x, grid_y, grid_z, cell_w, cell_h, width, height, cap, num_lights;
    float r_pad_column;
    float slice_near[HL_MAX_SLICES];
    float slice_far[HL_MAX_SLICES];
    float slice_r_pad[HL_MAX_SLICES];
};

}  // namespace hl

// ------------------------------------------------------------------ context
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct hl_external;

struct hl_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    DevBuf hot, cull, geom, bits, summary, raw, markq, shadeq, words, accum, vogel, rcirr, tilebin;
    bool vogel_ready = false;  // vogel holds the Vogel disk tables
    // this device's address of the no-alloc scratch array (hl_api.hip),
    // looked up once per device when the context is created
    void* noalloc = nullptr;
    uint32_t shade_parity = 0;  // which half of shadeq the next shading launch claims from
    // the lights the records in `hot` were last built from (stream order)
    const GpuLight* hot_src = nullptr;
    uint32_t hot_n = 0;
    // imported memory not yet released (hl_external_import); released with
    // the context
    std::vector<hl_external*> externals;
    // words: [0] overflow flags, [1] feature flags, [2..3] pair counter (u64)
};


namespace hl {

hl_status fail(hl_context* c, hl_status s, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
hl_status hip_fail(hl_context* c, hipError_t e, const char* what);
// Releases an imported allocation (hl_interop.hip): destroys the import and
// closes the library's descriptor.  The caller has synchronised the stream.
hl_status release_external(hl_external* e);
// Grows a scratch buffer (never shrinks).  zero: a new allocation is cleared
// (for buffers the kernels keep zero between calls).
hl_status ensure(hl_context* c, DevBuf& b, size_t bytes, bool zero = false);
hl_status check_gbuffer(hl_context* c, const hl_gbuffer* g, int* set);
hl_status check_cluster_params(hl_context* c, const hl_cluster_params* p);
ClusterArgs make_cluster_args(const hl_cluster_params* p, uint32_t num_lights);
uint32_t* ctx_words(hl_context* c);
hl_status build_shade_args(hl_context* c, const GpuCameraUniforms* camera_d,
                           const DeferredGlobals* g, const hl_gbuffer* gb, uint32_t num_lights,
                           const hl_light_grid* grid, ShadeArgs* a, int* set, bool need_debug0);

#define HL_HIP(c, expr)                                        \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return hl::hip_fail((c), _e, #expr); \
    } while (0)

// launchers (hl_cull.hip)
hipError_t launch_prep(hipStream_t s, const GpuLight* lights, const uint32_t* subset, uint32_t n,
                       const GpuCameraUniforms* cam, HotLight* hot, CullLight* cull,
                       uint32_t* flags);
// The adjoint's set-up: launch_prep's shading records (no subset, no
// culling records) with acc[0, acc_n) and *loss (when not null) cleared.
hipError_t launch_prep_zero(hipStream_t s, const GpuLight* lights, uint32_t n, const GpuCameraUniforms* cam,
                            HotLight* hot, uint32_t* flags, double* acc, size_t acc_n, double* loss);
hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n);
hipError_t launch_cull_tiles(hipStream_t s, const GpuCameraUniforms* cam, const LightCullParams& p,
                             const CullLight* cl, uint32_t* lists, uint32_t* counts);
// The tile cull over candidates: each light's candidate tiles (a conservative
// rectangle from necessary conditions of the plane test) are binned into
// super-tiles of kSuperTile x kSuperTile tiles; the super-tiles' ascending
// candidate lists come from the cluster count and fill kernels; each tile then
// tests its super-tile's candidates in order with the exact plane test.  Same
// lists as launch_cull_tiles when the tile grid is the screen's
// (num_tiles = ceil(size / 16)).  bits and summary: as the cluster cull's
// (zero before, zero after), over nsup = super-tiles cells.
constexpr uint32_t kSuperTile = 4;
struct TileBinScratch {
    uint32_t* bits;     // nsup * ceil(L/32) words
    uint32_t* summary;  // nsup * ceil(L/1024) words
    uint32_t* raw;      // cluster_raw_bytes(nsup)
    uint32_t* offsets;  // nsup
    uint32_t* counts;   // nsup
    uint32_t* pool;     // pool_cap
    uint32_t pool_cap;  // nsup * L (no truncation)
    uint32_t* flags;    // the fill's overflow word (never set: no cap, no truncation)
    float* planes;      // 12 floats per tile: its four planes
};
inline uint32_t tile_bin_supers(const LightCullParams& p) {
    return ((p.num_tiles_x + kSuperTile - 1u) / kSuperTile) * ((p.num_tiles_y + kSuperTile - 1u) / kSuperTile);
}
hipError_t launch_cull_tiles_binned(hipStream_t s, const GpuCameraUniforms* cam, const LightCullParams& p,
                                    const CullLight* cl, uint32_t* lists, uint32_t* counts,
                                    const TileBinScratch& w);
// Scratch of the clustered cull.  bits (ncell * ceil(L/32) words) and summary
// (ncell * ceil(L/1024) words) must be zero before the first call; every call
// leaves them zero again.
// Work queue of the cluster mark.  A light's candidate cells are enumerated
// in chunks of 64; the wave of a light marks its first kMarkInline chunks and
// queues the rest, which a second launch spreads over all CUs (a light close
// to the camera can have thousands of candidate cells).  The queue is split
// into kMarkQueues sub-queues (light i uses sub-queue i % kMarkQueues), each
// with its own counter on its own 128-B line, so that the claims of
// thousands of lights do not serialize on one address.
constexpr uint32_t kMarkInline = 2;
constexpr uint32_t kMarkQueues = 64;
constexpr uint32_t kMarkCountStride = 32;  // u32 words between counters
struct MarkQueue {
    uint32_t* count;  // kMarkQueues counters, reset by the geometry kernel (may exceed seg)
    uint4* bounds;    // per light: x0, y0, s0 << 16 | nz, nx << 16 | ny
    uint2* entries;   // kMarkQueues segments of seg entries (light, chunk)
    uint32_t seg;
};
inline uint32_t mark_queue_seg(uint32_t num_lights) {
    return (4u * num_lights + 4096u + kMarkQueues - 1u) / kMarkQueues;
}
inline size_t mark_queue_bytes(uint32_t num_lights) {
    return 4 * (size_t)kMarkQueues * kMarkCountStride + 16 * (size_t)num_lights +
           8 * (size_t)kMarkQueues * mark_queue_seg(num_lights);
}
inline MarkQueue mark_queue_at(void* base, uint32_t num_lights) {
    MarkQueue q;
    const size_t counts = 4 * (size_t)kMarkQueues * kMarkCountStride;
    q.count = (uint32_t*)base;
    q.bounds = (uint4*)((uint8_t*)base + counts);
    q.entries = (uint2*)((uint8_t*)base + counts + 16 * (size_t)num_lights);
    q.seg = mark_queue_seg(num_lights);
    return q;
}

struct ClusterScratch {
    void* geom;          // cluster_geom_bytes(columns, cells)
    uint32_t* bits;
    uint32_t* summary;
    uint32_t* raw;       // cluster_raw_bytes(ncell)
    MarkQueue mq;        // mark_queue_bytes(lights)
};
size_t cluster_geom_bytes(uint32_t ncol, uint32_t ncell);
size_t cluster_geom_offset_boxes(uint32_t ncol);
inline size_t cluster_bits_words(uint32_t ncell, uint32_t num_lights) {
    return (size_t)ncell * ((num_lights + 31u) / 32u);
}
inline size_t cluster_summary_words(uint32_t ncell, uint32_t num_lights) {
    return (size_t)ncell * (((num_lights + 31u) / 32u + 31u) / 32u);
}
// Raw per-cluster counts, then (256-B aligned) one 64-bit sum of capped
// counts per scan tile of kScanTile clusters.
constexpr uint32_t kScanTile = 1024;
__host__ __device__ inline uint32_t cluster_scan_tiles_dev(uint32_t ncell) {
    return (ncell + kScanTile - 1u) / kScanTile;
}
inline uint32_t cluster_scan_tiles(uint32_t ncell) { return cluster_scan_tiles_dev(ncell); }
__host__ __device__ inline size_t cluster_tile_sums_offset(uint32_t ncell) {
    return (4 * (size_t)ncell + 255) & ~(size_t)255;
}
inline size_t cluster_raw_bytes(uint32_t ncell) {
    return cluster_tile_sums_offset(ncell) + 8 * (size_t)cluster_scan_tiles(ncell);
}

// One view of a batched cull (device memory): its cluster arguments, camera,
// light subset, and its own per-view buffers.
// Each view's cell range in a batched cull starts on a multiple of this (the
// cells of one block of the geometry kernel), so no block spans two views.
constexpr uint32_t kBatchCellAlign = 256;
struct BatchCullView {
    ClusterArgs ca;
    const GpuCameraUniforms* cam;
    const uint32_t* subset;  // pool indices, ascending (nullptr: the first ca.num_lights)
    HotLight* hot;
    CullLight* cull;
    void* geom;              // cluster_geom_bytes(columns, cells)
    uint32_t* bits;
    uint32_t* summary;
    uint32_t* raw;           // cluster_raw_bytes(ncell)
    uint32_t* offsets;
    uint32_t* counts;
    uint32_t* pool;
    uint32_t* flags;
    uint32_t pool_cap, W, SW, ncell;
};
hipError_t launch_batch_cull(hipStream_t s, const GpuLight* pool, const BatchCullView* views,
                             const uint32_t* light_base, const uint32_t* cell_base,
                             const uint32_t* tile_base, uint32_t nviews, uint32_t total_lights,
                             uint32_t total_cells, uint32_t total_tiles, const MarkQueue& mq);

// Prepares the lights' records (hot: shading, may be null; cl: culling; feature
// flags) and builds the lists.
hipError_t launch_cluster_cull(hipStream_t s, const ClusterArgs& a, const GpuCameraUniforms* cam,
                               const GpuLight* lights, HotLight* hot, CullLight* cl, uint32_t* feature_flags,
                               const ClusterScratch& w, uint32_t* offsets,
                               uint32_t* counts, uint32_t* pool, uint32_t pool_cap,
                               uint32_t* flags, uint32_t* flags2);

// launchers (hl_shade.hip)
// kModeShadowed: kModeForward with shadow_factor on the shadow-casting lights
// (ShadeArgs.shadow.layers != 0); kModeExtended: kModeShadowed plus the
// radiance-cascade indirect diffuse (ShadeArgs.rc_fmt != 0), kept apart so
// that the shadow path keeps its occupancy
enum ShadeMode : int {
    kModeForward = 0,
    kModeCount = 1,
    kModeBackward = 2,
    kModeFusedL2 = 3,
    kModeShadowed = 4,
    kModeExtended = 5,
    // flag on the forward-writing modes (forward, shadowed, extended): the
    // output target is RGBA16F (ShadeArgs::out_f16) instead of RGBA32F
    kModeOutF16 = 8
};
// The fused forward + L2 adjoint shares a wave tile's lights among the block's
// four waves on views of fewer than kFusedSplitBlocks block tiles: a 512^2
// view has 1024 block tiles, 4096 waves of their own for 1024 SIMDs, and the
// sharing gives it four times as many.  Larger views give each wave a tile of
// its own (no partial sums across waves, no block barriers).  The rule
// depends on the view's size only, so every path shading a view picks the
// same split, and the same sums.
constexpr uint32_t kFusedSplitBlocks = 4096;
__host__ __device__ inline bool fused_split(uint32_t num_blocks) { return num_blocks < kFusedSplitBlocks; }

// Work items of a shading launch over num_blocks block tiles: the backward
// mode, and the fused mode when split, share each wave tile's lights among a
// block's four waves, so their items are wave tiles (four per block tile).
__host__ __device__ inline uint32_t shade_items(int mode, uint32_t num_blocks) {
    return (mode == kModeBackward || (mode == kModeFusedL2 && fused_split(num_blocks))) ? 4u * num_blocks
                                                                                       : num_blocks;
}

// Claim counters of one persistent shading launch: kShadeQueues counters,
// kShadeQueueStride words apart, zeroed beforehand (launch_fill_u32).
constexpr uint32_t kShadeQueues = 64;
constexpr uint32_t kShadeQueueStride = 32;
constexpr uint32_t kShadeQueueWords = kShadeQueues * kShadeQueueStride;
// One view, persistent: queue = zeroed claim counters (kShadeQueueWords),
// queue_next = the counter set to zero for the next launch on the stream.
hipError_t launch_shade(hipStream_t s, int gbuf_set, int mode, const ShadeArgs& a, uint32_t* queue,
                        uint32_t* queue_next);
// block tiles of a view for a wavefront tile width (1 << tw_log2)
inline void set_tile_shape(ShadeArgs& a, uint32_t tw_log2) {
    const uint32_t bw = 4u << tw_log2, bh = 64u >> tw_log2;
    a.tw_log2 = tw_log2;
    a.block_tiles_x = (a.width + bw - 1) / bw;
    a.num_blocks = a.block_tiles_x * ((a.height + bh - 1) / bh);
}
// Packs surface attributes into the colour targets of gb (format set gbuf_set).
hipError_t launch_gbuffer_pack(hipStream_t s, int gbuf_set, const hl_surface& sf, const hl_gbuffer& gb);

// launchers (hl_forward.hip)
// Everything a forward-lit launch reads (passed by value); the list walk reads
// the grid and light fields by these names.
struct ForwardArgs {
    const HotLight* lights;
    uint32_t num_lights;
    const uint32_t* lists;
    const uint32_t* counts;
    const uint32_t* offsets;
    uint32_t lists_capacity;
    uint32_t grid_mode;               // HL_GRID_TILES / HL_GRID_CLUSTERS
    uint32_t grid_x, grid_y, grid_z;  // tiles: num_tiles_x, num_tiles_y, 1
    uint32_t cell_w, cell_h;          // clusters: pixels per cell
    uint32_t tile_stride_x;           // tiles: ForwardLitGlobals.num_tiles_x
    const GpuCameraUniforms* cam;
    const hl_fragment* frags;
    uint32_t num_frags;
    float amb[3];                     // ambient_color.rgb * ambient_intensity
    float* out;                       // RGBA32F per fragment
    unsigned long long* pair_counter; // set: count pairs instead of shading
    float depth_split[HL_MAX_SLICES + 1];
};
// Shading records of the lights with the positions and directions of their
// transforms (entity == nullptr: the lights' own).
hipError_t launch_forward_prep(hipStream_t s, const GpuLight* lights, uint32_t n, const uint32_t* entity,
                               const hl_transform* transforms, uint32_t num_transforms, HotLight* hot);
hipError_t launch_forward_lit(hipStream_t s, const ForwardArgs& a);

// spans: nspans (first item, view) pairs in item order, then (num_items, 0).
// queue: kShadeQueueWords claim counters, zero before the first launch; each
// launch leaves them zero (its last block out clears them).  one: the
// arguments of the only view when nspans == 1 (then passed as a kernel
// argument), else null.
hipError_t launch_shade_batch(hipStream_t s, int gbuf_set, int mode, const BatchShadeView* views,
                              const uint2* spans, uint32_t nspans, uint32_t num_items, uint32_t* queue,
                              uint32_t grid_blocks, const ShadeArgs* one);
// Blocks of the batched shading kernel of (gbuf_set, mode) resident per CU.
uint32_t shade_batch_blocks_per_cu(int gbuf_set, int mode);
// Per-light FP64 gradient accumulators: kGradSlots slot-major copies of 4
// doubles (dpos.xyz, dI) per light, summed by launch_light_grads_to_f32.
constexpr uint32_t kGradSlots = 8;
__host__ __device__ inline size_t grad_accum_doubles(uint32_t num_lights) {
    return (size_t)kGradSlots * 4 * (num_lights ? num_lights : 1);
}
// The Vogel disk tables of the shadowed shading (kVogelFloats floats).
constexpr size_t kVogelTableFloats = 2 * 256 + (size_t)257 * 256;
hipError_t launch_vogel_table(hipStream_t s, float* t);
// a.rc_irr[pixel] = sample_rc_irradiance at every non-sky pixel (G-buffer format set gbuf_set)
hipError_t launch_rc_irradiance(hipStream_t s, int gbuf_set, const ShadeArgs& a);
hipError_t launch_light_grads_to_f32(hipStream_t s, const double* acc, uint32_t n,
                                     const uint32_t* subset, float* d_pos, float* d_int);
// The adjoint views of a batched plan (device records built with the plan):
// one launch clears every view's accumulators (grad_accum_doubles(n)), loss
// and pool-sized FP32 gradient outputs ahead of the shading, one sums every
// view's accumulators into its outputs after it.  max_n: the largest n;
// max_words: the largest 3 * pool or 2 * grad_accum_doubles(n).
struct GradView {
    double* acc;
    double* loss;
    float* d_pos;
    float* d_int;
    const uint32_t* subset;
    uint32_t n, pool;
};
hipError_t launch_grad_views_zero(hipStream_t s, const GradView* views, uint32_t nv, size_t max_words);
hipError_t launch_grad_views_to_f32(hipStream_t s, const GradView* views, uint32_t nv, uint32_t max_n);

}  // namespace hl
