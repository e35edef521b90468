/*
 * helio_hip/types.h -- byte-exact C mirrors of the libhelio GPU records that the
 * deferred-lighting hot path reads.  Every struct is asserted against the size
 * and offsets the Rust side asserts, so a buffer written by the Rust renderer
 * can be handed to this library unchanged.
 *
 * Sources (Rust, repr(C)):
 *   GpuLight            crates/libhelio/src/light.rs:60-106   (size assert :111, 16-multiple :115)
 *   GpuCameraUniforms   crates/libhelio/src/camera.rs:12-31   (filled by GpuCameraUniforms::new :33-58)
 *   GpuMaterial         crates/libhelio/src/material.rs:231-254
 *   LightCullParams     crates/passes/3d/helio-pass-light-cull/src/lib.rs:24-35
 *                       (WGSL mirror light_cull.wgsl:34-43)
 *   DeferredGlobals     crates/passes/3d/helio-pass-deferred-light/src/lib.rs:15-52
 *                       (WGSL mirror deferred_lighting.wgsl:33-67)
 *
 * Matrices are column-major f32[16] (glam Mat4::to_cols_array): element
 * (row r, column c) lives at index c*4 + r.
 */
#ifndef HELIO_HIP_TYPES_H
#define HELIO_HIP_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LightType discriminant, crates/libhelio/src/light.rs:8-13.  Area lights take
 * the point-light branch of pbr_direct_light (deferred_lighting.wgsl:724-775). */
enum {
    HL_LIGHT_DIRECTIONAL = 0,
    HL_LIGHT_POINT = 1,
    HL_LIGHT_SPOT = 2,
    HL_LIGHT_AREA = 3
};

/* shadow_index value meaning "casts no shadow" (deferred_lighting.wgsl:459). */
#define HL_NO_SHADOW 0xFFFFFFFFu

/* Reference tile grid constants, light_cull.wgsl:17-18 and
 * helio-pass-light-cull/src/lib.rs:17-18. */
#define HL_TILE_SIZE 16u
#define HL_MAX_LIGHTS_PER_TILE 64u

typedef struct GpuLight {
    float position_range[4];   /*   0: xyz world position, w range            */
    float direction_outer[4];  /*  16: xyz direction, w cos(outer cone angle)  */
    float color_intensity[4];  /*  32: xyz linear RGB, w intensity             */
    uint32_t shadow_index;     /*  48: HL_NO_SHADOW = no shadow                */
    uint32_t light_type;       /*  52: HL_LIGHT_*                              */
    float inner_angle;         /*  56: cos(inner cone angle)                   */
    uint32_t _pad;             /*  60                                          */
    uint32_t god_rays_enabled; /*  64                                          */
    float god_rays_density;    /*  68                                          */
    float god_rays_weight;     /*  72                                          */
    float god_rays_decay;      /*  76                                          */
    float god_rays_exposure;   /*  80                                          */
    uint32_t flare_enabled;    /*  84                                          */
    uint32_t flare_type;       /*  88                                          */
    float flare_intensity;     /*  92                                          */
    float flare_scale;         /*  96                                          */
    float flare_tint_r;        /* 100                                          */
    float flare_tint_g;        /* 104                                          */
    float flare_tint_b;        /* 108                                          */
    int32_t ies_profile_index; /* 112: -1 = none                               */
    int32_t light_function_index; /* 116: -1 = none                            */
    float ies_angle_scale;     /* 120                                          */
    float ies_angle_offset;    /* 124                                          */
} GpuLight;

typedef struct GpuCameraUniforms {
    float view[16];            /*   0 */
    float proj[16];            /*  64 */
    float view_proj[16];       /* 128 */
    float inv_view_proj[16];   /* 192 */
    float position_near[4];    /* 256: xyz camera position, w near plane */
    float forward_far[4];      /* 272: xyz forward, w far plane          */
    float jitter_frame[4];     /* 288                                    */
    float prev_view_proj[16];  /* 304                                    */
} GpuCameraUniforms;

typedef struct LightCullParams {
    uint32_t num_tiles_x;
    uint32_t num_tiles_y;
    uint32_t num_lights;
    uint32_t screen_width;
    uint32_t screen_height;
    uint32_t _pad0, _pad1, _pad2;
} LightCullParams;

typedef struct DeferredGlobals {
    uint32_t frame;                    /*   0 */
    float delta_time;                  /*   4 */
    uint32_t light_count;              /*   8 */
    float ambient_intensity;           /*  12 */
    float ambient_color[4];            /*  16 */
    float rc_world_min[4];             /*  32 */
    float rc_world_max[4];             /*  48 */
    float csm_splits[4];               /*  64 */
    uint32_t debug_mode;               /*  80 */
    uint32_t has_rc_gi;                /*  84 */
    uint32_t num_tiles_x;              /*  88 */
    uint32_t reflection_capture_count; /*  92 */
    uint32_t enable_reflections;       /*  96 */
    uint32_t enable_env_reflections;   /* 100 */
    uint32_t _pad[2];                  /* 104 */
} DeferredGlobals;

typedef struct GpuMaterial {
    float base_color[4];         /*  0 */
    float emissive[4];           /* 16 */
    float roughness_metallic[4]; /* 32: x roughness, y metallic, z ior, w specular */
    uint32_t tex_base_color;     /* 48 */
    uint32_t tex_normal;         /* 52 */
    uint32_t tex_roughness;      /* 56 */
    uint32_t tex_emissive;       /* 60 */
    uint32_t tex_occlusion;      /* 64 */
    uint32_t workflow;           /* 68 */
    uint32_t flags;              /* 72 */
    uint32_t material_class;     /* 76 */
    float class_params[4];       /* 80 */
} GpuMaterial;

/* ForwardLitGlobals, the forward-lit pass's uniform
 * (crates/passes/3d/helio-pass-forward-lit/src/lib.rs:10-22,
 * shaders/forward_lit.wgsl:17-27). */
typedef struct ForwardLitGlobals {
    uint32_t frame;                /*  0 */
    float delta_time;              /*  4 */
    uint32_t light_count;          /*  8 */
    float ambient_intensity;       /* 12 */
    float ambient_color[4];        /* 16 */
    uint32_t num_tiles_x;          /* 32 */
    uint32_t num_tiles_y;          /* 36 */
    float screen_width;            /* 40 */
    float screen_height;           /* 44 */
} ForwardLitGlobals;

/* SceneDB's Transform component as the forward-lit shader binds it: three
 * packed [f32; 3] (forward_lit.wgsl:110-122), rotation in degrees
 * [pitch_x, yaw_y, roll_z]. */
typedef struct hl_transform {
    float position[3];             /*  0 */
    float rotation[3];             /* 12 */
    float scale[3];                /* 24 */
} hl_transform;

/* Per-cascade and global shadow settings (crates/libhelio/src/shadow.rs:32-63). */
typedef struct CascadeConfig {
    float split_distance;          /*  0 */
    float depth_bias;              /*  4 */
    float filter_radius;           /*  8: PCF radius in atlas texels */
    float pcss_light_size;         /* 12: 0 disables PCSS for the cascade */
} CascadeConfig;

typedef struct ShadowConfig {
    CascadeConfig cascades[4];     /*  0 */
    uint32_t enable_pcss;          /* 64 */
    uint32_t pcss_blocker_samples; /* 68 */
    uint32_t pcss_filter_samples;  /* 72 */
    uint32_t pcf_sample_count;     /* 76 */
} ShadowConfig;

/* One reflection capture (crates/libhelio/src/reflection.rs:64-80; WGSL
 * deferred_lighting.wgsl:189-198), 112 bytes.  world_to_local is column-major
 * (WGSL mat4x4: column c at [4c .. 4c+3]). */
enum { HL_CAPTURE_SHAPE_SPHERE = 0, HL_CAPTURE_SHAPE_BOX = 1 };
typedef struct GpuReflectionCapture {
    float position_radius[4];    /*  0: xyz world position, w influence radius   */
    float extents_transition[4]; /* 16: xyz local half-extents (box), w transition */
    float world_to_local[16];    /* 32: box parallax; identity for spheres         */
    int32_t cubemap_index;       /* 96: cube of the array, -1 = none resident      */
    uint32_t shape;              /* 100: HL_CAPTURE_SHAPE_*                         */
    uint32_t mobility;           /* 104: 0 static, 1 dynamic (not read)            */
    float brightness;            /* 108                                             */
} GpuReflectionCapture;

#ifdef __cplusplus
#define HL_STATIC_ASSERT static_assert
#else
#define HL_STATIC_ASSERT _Static_assert
#endif

HL_STATIC_ASSERT(sizeof(GpuLight) == 128, "GpuLight must be 128 bytes (light.rs:111)");
HL_STATIC_ASSERT(sizeof(GpuLight) % 16 == 0, "GpuLight stride must be a 16-multiple (light.rs:115)");
HL_STATIC_ASSERT(offsetof(GpuLight, direction_outer) == 16, "GpuLight.direction_outer");
HL_STATIC_ASSERT(offsetof(GpuLight, color_intensity) == 32, "GpuLight.color_intensity");
HL_STATIC_ASSERT(offsetof(GpuLight, shadow_index) == 48, "GpuLight.shadow_index");
HL_STATIC_ASSERT(offsetof(GpuLight, light_type) == 52, "GpuLight.light_type");
HL_STATIC_ASSERT(offsetof(GpuLight, inner_angle) == 56, "GpuLight.inner_angle");
HL_STATIC_ASSERT(offsetof(GpuLight, god_rays_enabled) == 64, "GpuLight.god_rays_enabled");
HL_STATIC_ASSERT(offsetof(GpuLight, flare_enabled) == 84, "GpuLight.flare_enabled");
HL_STATIC_ASSERT(offsetof(GpuLight, flare_scale) == 96, "GpuLight.flare_scale");
HL_STATIC_ASSERT(offsetof(GpuLight, ies_profile_index) == 112, "GpuLight.ies_profile_index");
HL_STATIC_ASSERT(offsetof(GpuLight, light_function_index) == 116, "GpuLight.light_function_index");
HL_STATIC_ASSERT(offsetof(GpuLight, ies_angle_offset) == 124, "GpuLight.ies_angle_offset");

HL_STATIC_ASSERT(sizeof(GpuCameraUniforms) == 368, "GpuCameraUniforms must be 368 bytes");
HL_STATIC_ASSERT(offsetof(GpuCameraUniforms, proj) == 64, "camera.proj");
HL_STATIC_ASSERT(offsetof(GpuCameraUniforms, view_proj) == 128, "camera.view_proj");
HL_STATIC_ASSERT(offsetof(GpuCameraUniforms, inv_view_proj) == 192, "camera.inv_view_proj");
HL_STATIC_ASSERT(offsetof(GpuCameraUniforms, position_near) == 256, "camera.position_near");
HL_STATIC_ASSERT(offsetof(GpuCameraUniforms, forward_far) == 272, "camera.forward_far");
HL_STATIC_ASSERT(offsetof(GpuCameraUniforms, jitter_frame) == 288, "camera.jitter_frame");
HL_STATIC_ASSERT(offsetof(GpuCameraUniforms, prev_view_proj) == 304, "camera.prev_view_proj");

HL_STATIC_ASSERT(sizeof(LightCullParams) == 32, "LightCullParams must be 32 bytes");

HL_STATIC_ASSERT(sizeof(DeferredGlobals) == 112, "DeferredGlobals must be 112 bytes");
HL_STATIC_ASSERT(offsetof(DeferredGlobals, ambient_color) == 16, "globals.ambient_color");
HL_STATIC_ASSERT(offsetof(DeferredGlobals, csm_splits) == 64, "globals.csm_splits");
HL_STATIC_ASSERT(offsetof(DeferredGlobals, debug_mode) == 80, "globals.debug_mode");
HL_STATIC_ASSERT(offsetof(DeferredGlobals, num_tiles_x) == 88, "globals.num_tiles_x");
HL_STATIC_ASSERT(offsetof(DeferredGlobals, enable_env_reflections) == 100, "globals.enable_env_reflections");

HL_STATIC_ASSERT(sizeof(GpuMaterial) == 96, "GpuMaterial must be 96 bytes");
HL_STATIC_ASSERT(offsetof(GpuMaterial, roughness_metallic) == 32, "material.roughness_metallic");
HL_STATIC_ASSERT(offsetof(GpuMaterial, workflow) == 68, "material.workflow");
HL_STATIC_ASSERT(offsetof(GpuMaterial, class_params) == 80, "material.class_params");

HL_STATIC_ASSERT(sizeof(ForwardLitGlobals) == 48, "ForwardLitGlobals must be 48 bytes");
HL_STATIC_ASSERT(offsetof(ForwardLitGlobals, num_tiles_x) == 32, "forward globals.num_tiles_x");
HL_STATIC_ASSERT(sizeof(hl_transform) == 36, "Transform must be 36 bytes (three packed [f32; 3])");
HL_STATIC_ASSERT(sizeof(CascadeConfig) == 16, "CascadeConfig must be 16 bytes");
HL_STATIC_ASSERT(sizeof(ShadowConfig) == 80, "ShadowConfig must be 80 bytes (shadow.rs:49-63)");
HL_STATIC_ASSERT(offsetof(ShadowConfig, pcf_sample_count) == 76, "ShadowConfig.pcf_sample_count");
HL_STATIC_ASSERT(offsetof(hl_transform, rotation) == 12, "transform.rotation");
HL_STATIC_ASSERT(sizeof(GpuReflectionCapture) == 112, "GpuReflectionCapture must be 112 bytes");
HL_STATIC_ASSERT(offsetof(GpuReflectionCapture, cubemap_index) == 96, "capture.cubemap_index");

#ifdef __cplusplus
}
#endif
#endif /* HELIO_HIP_TYPES_H */
