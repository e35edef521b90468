/*
 * synth.h -- deterministic synthetic scenes (libhelio_synth.so): camera,
 * GpuLight arrays and G-buffers of the workloads the library is measured and
 * checked on.  Host code.
 */
#ifndef HELIO_SYNTH_H
#define HELIO_SYNTH_H

#include <stdint.h>

#include "../../include/helio_hip/types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hl_synth_config {
    uint32_t width, height;
    uint32_t gbuf_format;          /* 0: RGBA8/RGBA16F, 1: all fp16, 2: all fp32 */
    uint32_t num_lights;
    float spot_fraction;           /* share of spot lights (exact count, random slots) */
    float spot_outer_min_deg, spot_outer_max_deg;
    float spot_inner_ratio;        /* inner half-angle = ratio * outer */
    float light_box_center[3], light_box_size[3];
    float range_min, range_max;
    float intensity_min, intensity_max;
    float color_min, color_max;
    float eye[3], target[3];
    float fov_y_deg, near_plane, far_plane;
    uint32_t num_boxes;
    float box_field_half;          /* boxes placed in [-h, h]^2 on the ground */
    float box_size_min, box_size_max, box_height_min, box_height_max;
    float depth_noise;             /* sigma of the view-depth noise */
    uint32_t _pad;
    uint64_t seed;
} hl_synth_config;

/* Each returns 0 on success, nonzero for an invalid configuration. */
int hl_synth_camera(const hl_synth_config* c, GpuCameraUniforms* out);
int hl_synth_lights(const hl_synth_config* c, GpuLight* out);
/* Planes at the configuration's formats (albedo/orm: 4, 8 or 16 B per texel;
 * normal/emissive: 8 or 16; depth: 4), rows of width texels, top row first. */
int hl_synth_gbuffer(const hl_synth_config* c, void* albedo, void* normal, void* orm, void* emissive,
                     float* depth);
/* n ascending indices of [0, pool), drawn without replacement. */
int hl_synth_subset(uint64_t seed, uint32_t view, uint32_t pool, uint32_t n, uint32_t* out);
uint32_t hl_synth_abi_size(void);

#ifdef __cplusplus
}
#endif
#endif /* HELIO_SYNTH_H */
