/*
 * helio_hip.h -- C ABI of libhelio_hip.so, the MI355X (gfx950) implementation of
 * Helio's deferred-lighting hot path: light assignment (reference 16-px tiles or
 * log-depth clusters) and per-pixel Cook-Torrance GGX shading over a G-buffer,
 * plus an adjoint of the shading and a batched multi-view launch.
 *
 * Contract (mirrors the pass contract of the reference, without wgpu types):
 *  - The caller owns every device buffer (allocate with hl_alloc or hipMalloc).
 *    A context owns only its scratch memory, plans and captured graphs.
 *  - Every call is asynchronous on the context stream (hl_set_stream) and
 *    returns an hl_status.  Nothing aborts or prints; the message of the last
 *    failure of a context is read with hl_last_error.  The reference returns
 *    helio_core::Error::InvalidPassConfig in the same places
 *    (crates/helio-core/src/error.rs:62, helio-pass-deferred-light/src/lib.rs:914-919).
 *  - A context is used by one host thread at a time (the reference runs passes
 *    single-threaded per frame, crates/helio-core/src/context.rs:15).  Distinct
 *    contexts may run concurrently on distinct streams.
 *  - Pointers named *_d are device pointers; everything else is host memory
 *    that is read during the call and not retained, except inside a batch plan
 *    (hl_batch_plan_create), which keeps the device pointers of its views.
 */
#ifndef HELIO_HIP_H
#define HELIO_HIP_H

#include "types.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define HL_API __declspec(dllexport)
#else
#define HL_API __attribute__((visibility("default")))
#endif

typedef int32_t hl_status;
#define HL_OK 0
#define HL_EINVAL (-1)       /* bad argument: null pointer, size, format, alignment   */
#define HL_EHIP (-2)         /* a HIP runtime call failed                              */
#define HL_ECAPACITY (-3)    /* cluster pool overflow: lists valid but truncated       */
#define HL_EUNSUPPORTED (-4) /* input needs a feature outside the hot path             */
#define HL_ENOMEM (-5)       /* device allocation failed                               */

typedef struct hl_context hl_context;
typedef struct hl_plan hl_plan;

/* ---------------------------------------------------------------- context */
HL_API hl_status hl_context_create(int device, hl_context** out);
HL_API hl_status hl_context_destroy(hl_context* ctx);
/* stream: a hipStream_t (NULL = the default stream).  Must be a default-priority stream. */
HL_API hl_status hl_set_stream(hl_context* ctx, void* stream);
HL_API void* hl_get_stream(const hl_context* ctx);
HL_API hl_status hl_synchronize(hl_context* ctx);
HL_API const char* hl_last_error(const hl_context* ctx);
HL_API const char* hl_status_string(hl_status s);
/* Library version string and the gfx target it was built for. */
HL_API const char* hl_version(void);

/* Device-memory helpers (a caller may use hipMalloc instead). */
HL_API hl_status hl_alloc(hl_context* ctx, size_t bytes, void** out_d);
HL_API hl_status hl_free(hl_context* ctx, void* ptr_d);
HL_API hl_status hl_memcpy_htod(hl_context* ctx, void* dst_d, const void* src, size_t bytes);
HL_API hl_status hl_memcpy_dtoh(hl_context* ctx, void* dst, const void* src_d, size_t bytes);
HL_API hl_status hl_memcpy_dtod(hl_context* ctx, void* dst_d, const void* src_d, size_t bytes);
HL_API hl_status hl_memset_d32(hl_context* ctx, void* dst_d, uint32_t value, size_t count);

/* Timing on the context stream (HIP events).  hl_event_elapsed_ms synchronizes on `end`. */
typedef struct hl_event hl_event;
HL_API hl_status hl_event_create(hl_context* ctx, hl_event** out);
HL_API hl_status hl_event_destroy(hl_event* ev);
HL_API hl_status hl_event_record(hl_context* ctx, hl_event* ev);
HL_API hl_status hl_event_elapsed_ms(hl_event* start, hl_event* end, float* ms);

/* ---------------------------------------------------------------- interop */
/* Memory of another API on the same GPU as device pointers for the *_d
 * arguments.  Replaces the wgpu resources the reference's passes bind: the
 * frame slots tile_light_lists, tile_light_counts and pre_aa
 * (crates/libhelio/src/frame.rs:228-234), handed over at
 * RenderGraph::replace_pass_at (crates/helio-core/src/graph/execution.rs:193).
 * A Vulkan VkDeviceMemory allocated exportable gives the descriptor with
 * vkGetMemoryFdKHR (VK_EXTERNAL_MEMORY_HANDLE_TYPE_OPAQUE_FD_BIT or
 * _DMA_BUF_BIT_EXT); a HIP allocation made with hipMemCreate and exported with
 * hipMemExportToShareableHandle (POSIX descriptor) gives one too.
 *  - hl_external_import: the caller keeps fd (the library imports and later
 *    closes a duplicate).  size: bytes to import, at most the descriptor's
 *    (a dma-buf's size is checked; a descriptor without one is refused).
 *  - hl_external_map: [offset, offset + size) of the import as a device
 *    pointer, valid until hl_external_release; offset a multiple of
 *    HL_EXTERNAL_MAP_ALIGN.  Linear memory only: an optimally tiled image is
 *    copied to or from an exported buffer on the other API's queue.
 *  - hl_external_release: waits for the context stream, then releases the
 *    import and every pointer mapped from it.  hl_context_destroy releases
 *    what is left.
 *  - Ordering between the two APIs is the caller's: wait for the other
 *    API's writes (its fence) before the calls that read the memory, and for
 *    the context stream (hl_synchronize) before the other API reads what the
 *    calls wrote.  HIP's external semaphores are not available on Linux. */
typedef struct hl_external hl_external;
#define HL_EXTERNAL_FD 1u            /* an opaque-fd or dma-buf file descriptor */
#define HL_EXTERNAL_MAP_ALIGN 256u   /* wgpu's minimum storage-buffer offset alignment */
HL_API hl_status hl_external_import(hl_context* ctx, int fd, uint64_t size, uint32_t kind, hl_external** out);
HL_API hl_status hl_external_map(hl_external* ext, uint64_t offset, uint64_t size, void** out_d);
HL_API hl_status hl_external_release(hl_external* ext);

/* --------------------------------------------------------------- textures */
/* Pixel formats the G-buffer attachments may use.  The reference binds
 * albedo/orm as Rgba8Unorm and normal/emissive as Rgba16Float
 * (crates/passes/3d/helio-pass-gbuffer/src/lib.rs:240-283) and its bind group
 * accepts any non-filterable float format, so fp16 and fp32 G-buffers bind
 * unchanged (helio-pass-deferred-light/src/lib.rs:203-236). */
enum {
    HL_FMT_RGBA8_UNORM = 1,
    HL_FMT_RGBA16_FLOAT = 2,
    HL_FMT_RGBA32_FLOAT = 3,
    HL_FMT_R32_FLOAT = 4,   /* depth (Depth32Float, standard Z in [0,1]) */
    HL_FMT_RG16_FLOAT = 5   /* lightmap UV */
};

typedef struct hl_texture {
    const void* data_d;    /* first texel of row 0 (top row) */
    uint32_t pitch_bytes;  /* bytes between rows; >= width * texel size, multiple of 4 */
    uint32_t format;       /* HL_FMT_* */
} hl_texture;

/* The G-buffer the deferred pass reads (libhelio GBufferViews,
 * crates/libhelio/src/frame.rs:29-38, plus depth).
 * Supported format sets (albedo+orm / normal+emissive):
 *   RGBA8_UNORM / RGBA16_FLOAT   (the reference formats)
 *   RGBA16_FLOAT / RGBA16_FLOAT  (fp16 G-buffer)
 *   RGBA32_FLOAT / RGBA32_FLOAT  (fp32 G-buffer)
 * Depth is always R32_FLOAT.
 * lightmap_uv.data_d must be NULL: every pixel then carries the (-1,-1)
 * "no lightmap" sentinel the G-buffer pass writes for unbaked instances
 * (deferred_lighting.wgsl:1036-1038).  Screen-space AO, SSS and anisotropy
 * inputs are the reference's fallbacks (AO 1, flags 0). */
typedef struct hl_gbuffer {
    uint32_t width, height;
    hl_texture albedo;      /* rgb albedo, a alpha                 */
    hl_texture normal;      /* xyz world normal, a F0.r            */
    hl_texture orm;         /* r ao, g roughness, b metallic, a F0.g */
    hl_texture emissive;    /* rgb emissive, a F0.b                */
    hl_texture depth;       /* NDC depth; >= 1.0 is sky            */
    hl_texture lightmap_uv; /* must be NULL (see above)            */
} hl_gbuffer;

/* ------------------------------------------------------------- light grid */
#define HL_MAX_SLICES 64

enum { HL_GRID_TILES = 1, HL_GRID_CLUSTERS = 2 };

/* Log-depth froxel grid.  Cell (x, y, z) covers pixels
 * [x*cell_w, (x+1)*cell_w) x [y*cell_h, (y+1)*cell_h) and NDC depths
 * [depth_split[z], depth_split[z+1]); its index is x + grid_x*(y + grid_y*z).
 * Slice z spans view distances near*(far/near)^(z/grid_z) .. ^( (z+1)/grid_z ).
 * Fill with hl_cluster_params_init; the same bytes must reach the cull and the
 * shading calls (both the HIP path and any CPU checker read the thresholds
 * from here, so pixel-to-cluster assignment is identical everywhere). */
typedef struct hl_cluster_params {
    uint32_t grid_x, grid_y, grid_z;
    uint32_t cell_w, cell_h;      /* pixels per cell */
    uint32_t width, height;       /* screen size */
    uint32_t cap;                 /* indices kept per cluster, lowest first; 0 = unbounded */
    float near_plane, far_plane;
    float r_pad_column;           /* max of slice_r_pad (informational: the cull reads slice_r_pad) */
    float reserved;
    float depth_split[HL_MAX_SLICES + 1];  /* NDC depth where slice k begins */
    float slice_near[HL_MAX_SLICES];       /* padded view-distance range of slice k */
    float slice_far[HL_MAX_SLICES];
    float slice_r_pad[HL_MAX_SLICES];      /* radius padding of slice k (reconstruction error bound) */
    /* slice_near, slice_far and slice_r_pad must be finite and non-decreasing
     * over the grid_z slices, the paddings non-negative (HL_EINVAL otherwise) */
} hl_cluster_params;

/* Host-side helper: compute the slice thresholds from the camera (a HOST copy
 * of GpuCameraUniforms).  Returns HL_EUNSUPPORTED for a projection that is not
 * a standard-Z perspective (proj[2][3] must be -1). */
HL_API hl_status hl_cluster_params_init(const GpuCameraUniforms* camera, uint32_t width,
                                        uint32_t height, uint32_t grid_x, uint32_t grid_y,
                                        uint32_t grid_z, uint32_t cap, hl_cluster_params* out);

/* The per-cell light lists the shading pass consumes.  Tile mode is the
 * reference's tile_light_lists / tile_light_counts pair
 * (crates/libhelio/src/frame.rs:230-234, 482-491); cluster mode is a compact
 * offsets/counts/pool triple. */
typedef struct hl_light_grid {
    uint32_t mode;                 /* HL_GRID_TILES or HL_GRID_CLUSTERS          */
    uint32_t num_tiles_x;          /* tiles: ceil(width/16)                       */
    uint32_t num_tiles_y;          /* tiles: ceil(height/16)                      */
    const uint32_t* lists_d;       /* tiles: [tile*64 + i]; clusters: pool        */
    const uint32_t* counts_d;      /* per cell                                    */
    const uint32_t* offsets_d;     /* clusters: per-cell start in the pool        */
    const hl_cluster_params* clusters; /* clusters: host copy of the params       */
    uint32_t lists_capacity;       /* entries readable at lists_d (tiles: 0 = tiles*64);
                                      list reads beyond it are treated as empty */
    uint32_t reserved;
} hl_light_grid;

/* --------------------------------------------------------------- culling */

/* LightCull, reference tile mode.  Replaces the dispatch in
 * LightCullPass::execute (crates/passes/3d/helio-pass-light-cull/src/lib.rs:284-362)
 * running light_cull.wgsl::main (shaders/light_cull.wgsl:123-180).
 * camera_d: the camera storage buffer (only element 0 is read).
 * params: LightCullParams as written by LightCullPass::prepare (lib.rs:264-282).
 * Outputs: lists_d[tile*64 + i] for i < counts_d[tile] hold the indices of the
 * lights whose sphere touches the tile's four lateral planes (directional
 * lights always), in ascending order, at most 64 (the lowest 64).
 * num_lights == 0 clears both outputs like lib.rs:285-292.
 * Bit-exact with the FP32 restatement of light_cull.wgsl. */
HL_API hl_status hl_light_cull_tiles(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                     const LightCullParams* params, const GpuLight* lights_d,
                                     uint32_t* lists_d, uint32_t* counts_d);

/* LightCull, clustered mode (log-depth froxels; no counterpart in the reference,
 * which only has tiles -- the north-star extension of light_cull.wgsl).
 * A light is listed in a cluster when
 *   directional: always;
 *   otherwise: its sphere (view-space centre, radius + padding) passes the four
 *   lateral planes of the cluster's screen cell AND overlaps the cluster's
 *   padded view-space AABB; spot lights must also pass a cone test against the
 *   AABB's bounding sphere (cone half-angle acos(direction_outer.w)).
 * Outputs: counts_d[c] = min(hits, cap) (cap 0 = unbounded); offsets_d[c] =
 * exclusive prefix sum of those counts in cluster order; pool_d[offsets[c] + i]
 * the ascending indices.  If the total exceeds pool_cap, clusters are truncated
 * (counts_d[c] = min(count, pool_cap - offsets[c]), offsets clamped to pool_cap)
 * so that nothing is written past pool_cap, and *overflow_d gets
 * HL_OVERFLOW_POOL; a cluster that hit its cap sets HL_OVERFLOW_CAP.  The call
 * still returns HL_OK (the word lives on the device; read it with
 * hl_query_overflow).
 * overflow_d may be NULL (the context's own word is used). */
HL_API hl_status hl_light_cull_clusters(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                        const hl_cluster_params* params,
                                        const GpuLight* lights_d, uint32_t num_lights,
                                        uint32_t* offsets_d, uint32_t* counts_d,
                                        uint32_t* pool_d, uint32_t pool_cap,
                                        uint32_t* overflow_d);

#define HL_OVERFLOW_CAP 1u
#define HL_OVERFLOW_POOL 2u

/* Reads back the overflow word of the last hl_light_cull_clusters call on
 * this context (synchronizes the stream).  0 = no truncation. */
HL_API hl_status hl_query_overflow(hl_context* ctx, uint32_t* flags);

/* --------------------------------------------------------- G-buffer pack */
/* Per-pixel surface attributes as the G-buffer pass's fragment shader has them
 * before it writes its targets (crates/passes/3d/helio-pass-gbuffer/shaders/
 * gbuffer.wgsl:480-493): RGBA32F planes of width x height, all with the same
 * row pitch. */
typedef struct hl_surface {
    uint32_t width, height;
    uint32_t pitch_bytes;            /* >= 16 * width, multiple of 16 */
    uint32_t reserved;               /* 0 */
    const float* albedo_d;           /* rgb albedo, a alpha                            */
    const float* normal_d;           /* xyz world normal (written as given), w ignored */
    const float* orm_d;              /* ao, roughness, metallic, w ignored              */
    const float* emissive_d;         /* rgb emissive, w ignored                         */
    const float* specular_f0_d;      /* rgb F0 of the specular workflow, w ignored;
                                        NULL = metallic workflow                        */
} hl_surface;

/* Writes the G-buffer the deferred pass reads, in the formats of `gbuffer`
 * (one of the format sets above; gbuffer's width/height must equal the
 * surface's; depth is not written):
 *   albedo   = (albedo.rgb, alpha)
 *   normal   = (N, F0.r)
 *   orm      = (ao, roughness, metallic, F0.g)
 *   emissive = (emissive.rgb, F0.b)             (gbuffer.wgsl:10-15, 488-491)
 * with F0 = clamp(mix(0.04, albedo, metallic), 0, 0.999) in the metallic
 * workflow (gbuffer.wgsl:324-328), or the given specular F0.  When
 * gbuffer->lightmap_uv.data_d is set (RG16_FLOAT) it is filled with the
 * (-1, -1) "no lightmap" sentinel.  Conversions: unorm8 = round-to-nearest-even
 * of clamp(v, 0, 1) * 255 (NaN -> 0); fp16 = IEEE round-to-nearest-even. */
HL_API hl_status hl_gbuffer_pack(hl_context* ctx, const hl_surface* surface, const hl_gbuffer* gbuffer);

/* ---------------------------------------------------------------- shading */
enum {
    HL_SHADE_CLEAR_SKY = 1u,        /* write (0,0,0,0) to sky pixels (LoadOp::Clear); default leaves them */
    HL_SHADE_LIGHTS_FROM_CULL = 2u, /* use the light records that this context's last light cull built
                                       (hl_light_cull_clusters / hl_light_cull_tiles over the same lights_d
                                       and num_lights, the buffer unchanged since in stream order) instead
                                       of building them again; HL_EINVAL when the last cull was over other
                                       lights.  The records depend on the lights only, not the camera. */
    HL_SHADE_OUT_RGBA16F = 4u       /* out_d is an RGBA16F target (the pass's pre_aa format,
                                       helio-pass-deferred-light/src/lib.rs:423-427) instead of RGBA32F:
                                       each channel of the FP32 result rounded to the nearest binary16
                                       (ties to even, overflow to inf); row pitch >= 8 * width, 8-byte
                                       aligned.  hl_deferred_forward, _shadowed and _ex only. */
};

/* DeferredLight fs_main over the G-buffer (deferred_lighting.wgsl:991-1259 with
 * the reference's fallbacks bound: no shadow casters, no IES/gobo textures, no
 * RC GI, no lightmaps, SSAO 1; fs_reflection then adds exactly 0,
 * deferred_lighting.wgsl:1265-1341).  Replaces DeferredLightPass::execute
 * (crates/passes/3d/helio-pass-deferred-light/src/lib.rs:913-1199).
 * globals: DeferredGlobals as written by DeferredLightPass::prepare (lib.rs:836-878);
 *   ambient_color/ambient_intensity/debug_mode/num_tiles_x are read.  Debug modes
 *   0, 1, 2, 4, 5, 20, 21 are supported; others return HL_EUNSUPPORTED.
 * out_d: RGBA32F target (pre_aa), row pitch out_pitch_bytes, or RGBA16F with
 *   HL_SHADE_OUT_RGBA16F.  Sky pixels (depth >= 1) are left untouched unless
 *   flags has HL_SHADE_CLEAR_SKY.
 * Lights with shadow_index != HL_NO_SHADOW or IES/gobo indices >= 0 are shaded
 *   as unshadowed, untextured lights (those inputs are outside the hot path).
 * Radiance is within 1e-4 relative of the FP32 restatement evaluating the same
 * lists. */
HL_API hl_status hl_deferred_forward(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                     const DeferredGlobals* globals, const hl_gbuffer* gbuffer,
                                     const GpuLight* lights_d, uint32_t num_lights,
                                     const hl_light_grid* grid, float* out_d,
                                     uint32_t out_pitch_bytes, uint32_t flags);

/* No counterpart in the reference: hl_deferred_forward for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Same arguments and the
 * same shading (bit-identical output); the light records and the claim
 * counters live in a fixed device array of the library
 * (HL_NOALLOC_SCRATCH_BYTES on each device, held until the process ends)
 * instead of context scratch, so the call allocates nothing.
 * Limits: num_lights <= HL_NOALLOC_MAX_LIGHTS (HL_EINVAL beyond); flags may
 * hold HL_SHADE_CLEAR_SKY only.  Every call on a device shares that device's
 * array: calls that overlap on one device (different streams or contexts,
 * not ordered by events) race and must be serialised by the caller. */
#define HL_NOALLOC_MAX_LIGHTS 65536u
#define HL_NOALLOC_SCRATCH_BYTES (16384u + 80u * HL_NOALLOC_MAX_LIGHTS)
HL_API hl_status hl_deferred_forward_noalloc(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                             const DeferredGlobals* globals, const hl_gbuffer* gbuffer,
                                             const GpuLight* lights_d, uint32_t num_lights,
                                             const hl_light_grid* grid, float* out_d,
                                             uint32_t out_pitch_bytes, uint32_t flags);

/* Shadow maps shadow_factor samples (deferred_lighting.wgsl:234-565): the
 * dynamic and the static shadow atlas (Depth32Float texture arrays bound at
 * :160-161), num_layers square layers of atlas_size texels (ATLAS_SIZE = 1024
 * at :236), each layer-major and tightly packed (texel (x, y) of layer l at
 * [(l * atlas_size + y) * atlas_size + x], row 0 = v 0), and one light
 * matrix per layer (column-major mat4, world -> light clip; shadow_matrices
 * at :167).  config is the ShadowConfig uniform (:139). */
typedef struct hl_shadow_maps {
    const float* atlas_d;
    const float* static_atlas_d;
    const float* matrices_d;
    uint32_t num_layers;
    uint32_t atlas_size;
    ShadowConfig config;
} hl_shadow_maps;

/* hl_deferred_forward with shadows: every term of a light i < 42
 * (MAX_SHADOW_LIGHTS, :20) whose shadow_index != HL_NO_SHADOW is multiplied by
 * shadow_factor (:452-565): normal-offset bias, the light's layer (point: base
 * + cube face of :253-263; spot: base; directional: base + cascade, blending
 * across the 10 % zones around globals->csm_splits with the camera distance),
 * then PCF (Vogel disk of pcf_sample_count taps scaled per cascade, :273-330)
 * or PCSS (blocker search + variable-radius PCF, :335-450) over both atlases,
 * lit = min(dynamic, static) per tap.  The comparison sampler (LessEqual,
 * bilinear, clamp to edge; helio-pass-shadow/src/lib.rs:365-375) is evaluated
 * with FP32 bilinear weights; sin and cos of the Vogel angles use a fixed
 * Cody-Waite reduction and Taylor polynomials, so the factor is reproducible
 * bit for bit.  A layer index >= num_layers samples nothing (factor 1).
 * shadows == NULL is hl_deferred_forward.  Same other arguments and results. */
HL_API hl_status hl_deferred_forward_shadowed(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                              const DeferredGlobals* globals, const hl_gbuffer* gbuffer,
                                              const GpuLight* lights_d, uint32_t num_lights,
                                              const hl_light_grid* grid, const hl_shadow_maps* shadows,
                                              float* out_d, uint32_t out_pitch_bytes, uint32_t flags);

/* hl_deferred_forward_shadowed plus the radiance-cascade indirect diffuse
 * (sample_rc_irradiance, deferred_lighting.wgsl:842-904, combined at
 * :1159-1163 and :1189-1205): when rc_cascade0 is given and
 * globals->has_rc_gi != 0, the ambient term becomes
 * mix(hemisphere, kD_ibl * rc_irr * albedo, clamp(4 |rc_irr|, 0, 1)) * ao with
 * kD_ibl from fresnel_schlick_roughness (pbr_eval.wgsl:63-66).  rc_cascade0 is
 * the RC probe atlas (bound at :168): RC_DIR_DIM * RC_PROBE_DIM = 64 texels
 * wide, 1024 rows, RGBA16_FLOAT or RGBA32_FLOAT, read texel by texel
 * (textureLoad) over the box globals->rc_world_min .. rc_world_max.  shadows
 * and rc_cascade0 may each be NULL. */
HL_API hl_status hl_deferred_forward_ex(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                        const DeferredGlobals* globals, const hl_gbuffer* gbuffer,
                                        const GpuLight* lights_d, uint32_t num_lights,
                                        const hl_light_grid* grid, const hl_shadow_maps* shadows,
                                        const hl_texture* rc_cascade0, float* out_d, uint32_t out_pitch_bytes,
                                        uint32_t flags);

/* ------------------------------------------------------------ reflections */
/* Pre-filtered reflection cube array (the baked cube array of
 * crates/helio-bake/src/data.rs:300-365, bound as env_cube at
 * deferred_lighting.wgsl:167).  Cube c holds mip_levels mips, mip m of
 * res_m = max(1, face_size >> m) texels per side; within a cube the mips follow
 * one another and within a mip the faces +X -X +Y -Y +Z -Z, each res_m rows of
 * res_m tightly packed texels.  Cube c starts cube_bytes(face_size, mip_levels,
 * format) * c bytes after data_d.  data_d NULL is the 1x1 black fallback cube
 * array (every sample 0). */
typedef struct hl_cube_array {
    const void* data_d;
    uint32_t face_size;   /* 1 .. 16384 */
    uint32_t mip_levels;  /* 1 .. 1 + log2(face_size) */
    uint32_t num_cubes;
    uint32_t format;      /* HL_FMT_RGBA32_FLOAT, HL_FMT_RGBA16_FLOAT or HL_FMT_RGBA8_UNORM */
} hl_cube_array;

/* Inputs of the reflection composite besides the G-buffer.  Every texture is
 * optional (data_d NULL binds the reference's black fallback); ssr and planar
 * are full-resolution RGBA16_FLOAT or RGBA32_FLOAT (a = hit confidence);
 * rc_cascade0 is the probe atlas of hl_deferred_forward_ex. */
typedef struct hl_reflection_inputs {
    hl_cube_array env;
    const GpuReflectionCapture* captures_d; /* min(globals->reflection_capture_count, 64) records */
    hl_texture ssr;
    hl_texture planar;
    hl_texture rc_cascade0;
} hl_reflection_inputs;

enum {
    HL_REFLECTION_STORE = 1u  /* write (contribution, 0) instead of adding it (the default, the
                                 pass's One + One blend over fs_main's output) */
};

/* The deferred pass's second draw, fs_reflection (deferred_lighting.wgsl:
 * 1266-1341; drawn at crates/passes/3d/helio-pass-deferred-light/src/lib.rs:
 * 1188-1197 with the One + One colour blend of :464-478, alpha kept): per
 * non-sky pixel,
 *   env = sample_reflection_environment(P, R, roughness * ENV_MAX_LOD) (:646-691:
 *         the captures front to back with capture_weight (:619-639) and box or
 *         sphere parallax (:577-615), layer 0 as the skylight for what is left)
 *         times F0 * env_brdf.x + env_brdf.y (pbr_eval.wgsl:70-76),
 *   then mixed toward ssr.rgb * F_ibl by ssr.a and planar.rgb * F_ibl by planar.a
 *   (enable_reflections), and toward sample_rc_specular (:929-983) by
 *   smoothstep(0.6, 0.9, roughness) * 0.4 where has_rc_gi and roughness > 0.6,
 *   times ao (the screen-space AO input is its fallback, 1);
 * out.rgb += that, or out = (that, 0) with HL_REFLECTION_STORE.  The cube
 * sampler (linear, linear mips, clamp to edge; data.rs:372-380) is evaluated
 * with FP32 bilinear and mip weights on the face that the direction's major
 * axis selects (no filtering across faces); the cube index is clamped to
 * num_cubes - 1.  debug_mode 30 and 31 write the SSR diagnostics (a, a, a, 1)
 * and (rgb, 1) (:1272-1277) in place (the debug pipeline has no blend, lib.rs:489-510); 1, 2, 4, 5, 10, 11, 20 and 21 add nothing
 * (:1280-1285).  Sky pixels (depth >= 1) are left untouched.  Same G-buffer
 * rules as hl_deferred_forward (lightmap_uv NULL, so the contribution is
 * multiplied by ao). */
HL_API hl_status hl_deferred_reflection(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                        const DeferredGlobals* globals, const hl_gbuffer* gbuffer,
                                        const hl_reflection_inputs* inputs, float* out_d,
                                        uint32_t out_pitch_bytes, uint32_t flags);

/* Counts Q, the (pixel, light) pairs the forward pass evaluates past the range
 * test (type != 0 implies dist <= range) -- the "pixel-light evals" of the
 * throughput metric.  Same inputs as hl_deferred_forward; writes *pairs (host). */
HL_API hl_status hl_count_pixel_light_pairs(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                            const DeferredGlobals* globals,
                                            const hl_gbuffer* gbuffer, const GpuLight* lights_d,
                                            uint32_t num_lights, const hl_light_grid* grid,
                                            uint64_t* pairs);
/* Q as above, and the (wavefront, light) visits of the forward walk: the
 * lights the shading launch evaluates for 64 pixels at once (per wavefront,
 * per distinct cell of its pixels, the cell's lights past the test against
 * the bounds of those pixels).  Q / (64 visits) is the lane utilisation of the
 * per-light evaluation.  A diagnostic with no reference counterpart. */
HL_API hl_status hl_shade_walk_stats(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                     const DeferredGlobals* globals, const hl_gbuffer* gbuffer,
                                     const GpuLight* lights_d, uint32_t num_lights, const hl_light_grid* grid,
                                     uint64_t* pairs, uint64_t* wave_light_visits);

/* ------------------------------------------------------------ forward-lit */
/* One fragment of a forward-shaded surface: what ForwardLit's fs_main holds
 * after its surface evaluation (forward_lit.wgsl:245-282, 440-456), 80 bytes.
 * pixel is the framebuffer pixel the fragment covers (x | y << 16), whose
 * 16-px tile (or cluster) selects the light list (forward_lit.wgsl:463-466). */
typedef struct hl_fragment {
    float world_pos[3];
    uint32_t pixel;          /* x | y << 16                                     */
    float normal[3];         /* N as the surface evaluation returns it (unit)   */
    float alpha;             /* albedo.a; <= 0.001 discards the fragment        */
    float albedo[3];
    float ao;
    float specular_f0[3];    /* F0 as resolved by the surface evaluation        */
    float roughness;
    float emissive[3];
    float metallic;
} hl_fragment;

/* ForwardLit fs_main over a batch of fragments (forward_lit.wgsl:440-493), the
 * second consumer of the light grid (crates/libhelio/src/frame.rs:382-383).
 * Replaces the fragment stage of ForwardLitPass::execute
 * (crates/passes/3d/helio-pass-forward-lit/src/lib.rs:408-615); rasterisation,
 * depth test and blending stay with the caller.
 * Light world positions and directions come from
 * transforms_d[light_entity_indices_d[i]] (position; direction = -Z rotated by
 * the YXZ Euler rotation, forward_lit.wgsl:337-372), as the shader reads them;
 * range, cone angles, colour and type from lights_d.  With
 * light_entity_indices_d == NULL (and transforms_d NULL) the lights' own
 * position_range.xyz / direction_outer.xyz are used.  A light whose entity
 * index is >= num_transforms contributes nothing.
 * globals: ForwardLitGlobals as ForwardLitPass::prepare writes it
 *   (lib.rs:382-406); ambient_color, ambient_intensity, num_tiles_x/y and
 *   screen_width/height are read.
 * grid: the light lists -- tiles (the reference's tile_light_lists /
 *   tile_light_counts, indexed with globals.num_tiles_x) or clusters (the
 *   fragment's slice from the NDC depth of its world position).
 * out_d: one RGBA32F value per fragment, in fragment order:
 *   (Lo + ambient + emissive, alpha); a discarded fragment (alpha <= 0.001)
 *   gets (0, 0, 0, 0).  A fragment whose pixel lies outside the grid is lit by
 *   no light (ambient and emissive only).
 * fragments_d and out_d are 16-byte aligned.  flags must be 0.
 * Fragments in screen-coherent order (as a rasteriser emits them, or sorted
 * by tile) share light lists within a wavefront; any order is correct.
 * Radiance is within 1e-4 relative of the FP32 restatement evaluating the same
 * lists. */
HL_API hl_status hl_forward_lit(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                const ForwardLitGlobals* globals, const hl_fragment* fragments_d,
                                uint32_t num_fragments, const GpuLight* lights_d, uint32_t num_lights,
                                const uint32_t* light_entity_indices_d,
                                const hl_transform* transforms_d, uint32_t num_transforms,
                                const hl_light_grid* grid, float* out_d, uint32_t flags);

/* Counts the (fragment, light) pairs hl_forward_lit evaluates past the range
 * test (kept fragments only) -- its "fragment-light evals".  Same inputs;
 * writes *pairs (host). */
HL_API hl_status hl_forward_lit_count_pairs(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                            const ForwardLitGlobals* globals,
                                            const hl_fragment* fragments_d, uint32_t num_fragments,
                                            const GpuLight* lights_d, uint32_t num_lights,
                                            const uint32_t* light_entity_indices_d,
                                            const hl_transform* transforms_d,
                                            uint32_t num_transforms, const hl_light_grid* grid,
                                            uint64_t* pairs);

/* ---------------------------------------------------------------- adjoint */
/* Per-pixel gradient planes: d_pixel_d holds 5 planes of width*height floats,
 * plane-major: dL/d albedo.r, .g, .b, dL/d roughness, dL/d metallic. */
#define HL_PIXEL_GRAD_PLANES 5

/* Adjoint of hl_deferred_forward for debug mode 0 with respect to per-pixel
 * albedo, roughness, metallic and per-light position and intensity.  The
 * differentiated forward resolves F0 from albedo and metallic the way the
 * G-buffer producer does (F0 = clamp(mix(0.04, albedo, metallic), 0, 0.999),
 * crates/passes/3d/helio-pass-gbuffer/shaders/gbuffer.wgsl:324-328) instead of
 * reading the stored F0 channels (an 8-bit orm.w stores a dielectric's 0.04 as
 * 10/255, 2 % lower); with an fp32 G-buffer written by that producer both
 * forwards agree up to the order of the sums (the fused forward adds four
 * partial sums per pixel: within 1e-6 relative).  The no-contribution cut-offs
 * (dist > range, NdL == 0, all(radiance < 0.002)) are treated as fixed.
 * dL_dout_d: RGBA32F per pixel (alpha ignored), pitch dout_pitch_bytes.
 * Per-light outputs are overwritten: d_light_pos_d[3*i + k], d_light_intensity_d[i].
 * The per-cell lists are re-derived per wave from the grid, never stored. */
HL_API hl_status hl_deferred_backward(hl_context* ctx, const GpuCameraUniforms* camera_d,
                                      const DeferredGlobals* globals, const hl_gbuffer* gbuffer,
                                      const GpuLight* lights_d, uint32_t num_lights,
                                      const hl_light_grid* grid, const float* dL_dout_d,
                                      uint32_t dout_pitch_bytes, float* d_pixel_d,
                                      float* d_light_pos_d, float* d_light_intensity_d);

/* Fused forward + per-pixel L2 loss against target_d + adjoint.  Writes the
 * forward radiance to out_d, the gradients as hl_deferred_backward does with
 * dL/dout = 2 (out - target), and the loss sum_p |out_p - target_p|^2 over the
 * rgb channels of non-sky pixels to *loss_d (a device double, overwritten). */
HL_API hl_status hl_deferred_forward_backward_l2(
    hl_context* ctx, const GpuCameraUniforms* camera_d, const DeferredGlobals* globals,
    const hl_gbuffer* gbuffer, const GpuLight* lights_d, uint32_t num_lights,
    const hl_light_grid* grid, const float* target_d, uint32_t target_pitch_bytes, float* out_d,
    uint32_t out_pitch_bytes, double* loss_d, float* d_pixel_d, float* d_light_pos_d,
    float* d_light_intensity_d);

/* ------------------------------------------------------ batched multi-view */
/* One view of a batch: a camera, a G-buffer, a subset of a shared light pool
 * and an output.  The reference renders each view with its own full graph
 * submit (crates/helio/src/renderer/render.rs:804-873); a plan culls all views
 * with one launch per cull stage and shades them with one persistent launch
 * per G-buffer format set and mode (forward / forward + L2 adjoint), the whole
 * frame capturable in a HIP graph. */
typedef struct hl_view_desc {
    const GpuCameraUniforms* camera_d;
    GpuCameraUniforms camera;          /* host copy of the same camera (for slice set-up) */
    DeferredGlobals globals;
    hl_gbuffer gbuffer;
    const uint32_t* light_indices_d;   /* ascending pool indices, or NULL = whole pool */
    uint32_t num_lights;               /* entries of light_indices_d (pool size if NULL) */
    float* out_d;
    uint32_t out_pitch_bytes;
    /* Optional fused L2 adjoint (all NULL = forward only). */
    const float* target_d;
    uint32_t target_pitch_bytes;
    double* loss_d;
    float* d_pixel_d;                  /* 5 planes of width*height              */
    float* d_light_pos_d;              /* pool-sized: 3 floats per pool light   */
    float* d_light_intensity_d;        /* pool-sized                            */
    /* Optional fixed cluster grid (from hl_cluster_params_init for this view's
     * camera and size; cap honoured).  NULL = the plan chooses the grid (the
     * autotuner, when enabled, searches it); set = only the wave tile is tuned. */
    const hl_cluster_params* clusters;
} hl_view_desc;

/* Limits of a plan: at most HL_BATCH_MAX_VIEWS views, each of any size within
 * the single-view limits (the 50-view set of 14.6 M pixels runs as one plan). */
#define HL_BATCH_MAX_VIEWS 255u

typedef struct hl_autotune_opts {
    uint32_t enable;         /* 0 = use the default shape for every view              */
    uint32_t max_trials;     /* candidates timed per view class (0 = all)             */
    const char* cache_path;  /* file the results are kept in (NULL = no file cache)   */
} hl_autotune_opts;

/* Chosen configuration of one view (after hl_batch_plan_create). */
typedef struct hl_view_config {
    uint32_t wave_tile_w, wave_tile_h;   /* pixels one wavefront shades           */
    uint32_t grid_x, grid_y, grid_z;     /* cluster grid                          */
    float est_ms;                        /* autotuner's measured time, 0 if none  */
} hl_view_config;

HL_API hl_status hl_batch_plan_create(hl_context* ctx, const GpuLight* pool_d, uint32_t pool_size,
                                      const hl_view_desc* views, uint32_t num_views,
                                      const hl_autotune_opts* opts, hl_plan** out);
/* Record the plan's whole frame (light gather, clustering, shading, adjoints)
 * into a HIP graph; hl_batch_launch then replays it.  Without a capture,
 * hl_batch_launch runs the same launches directly. */
HL_API hl_status hl_batch_capture(hl_plan* plan);
HL_API hl_status hl_batch_launch(hl_plan* plan);
HL_API hl_status hl_batch_get_config(const hl_plan* plan, uint32_t view, hl_view_config* cfg);
HL_API hl_status hl_plan_destroy(hl_plan* plan);

#ifdef __cplusplus
}
#endif
#endif /* HELIO_HIP_H */
