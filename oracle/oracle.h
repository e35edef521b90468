/*
 * oracle.h -- CPU oracle of the deferred-lighting hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; only the tests, the smoke check and the CPU-baseline leg of
 * the benchmark use it, and only as the checker.
 *
 * Two parts:
 *  - FP32 restatement of the reference's WGSL, one operation at a time, built
 *    with -ffp-contract=off and IEEE sqrt/division:
 *      or_light_cull_tiles    crates/passes/3d/helio-pass-light-cull/shaders/light_cull.wgsl:85-180
 *      or_shade               crates/passes/3d/helio-pass-deferred-light/shaders/deferred_lighting.wgsl:991-1259
 *                             (pbr_direct_light :699-820, pbr_eval.wgsl:6-66)
 *    plus the clustered cull the HIP library adds (or_light_cull_clusters), whose
 *    membership rule is defined in helio_hip.h and restated here.
 *  - FP64 differentiable forward, analytic adjoint and central finite differences
 *    (or64_*), all lights evaluated per pixel (no grid).
 *  - The G-buffer producer's format contract (or_gbuffer_pack, gbuffer_pack.c).
 *  - The forward-lit fragment shader (or_forward_lit, forward_lit.c), the
 *    second consumer of the light grid.
 *
 * WGSL leaves a few operation orders to the backend; this oracle fixes them:
 *    dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
 *    length(v)     = sqrt(dot(v,v));   normalize(v) = v / length(v)   (3 divisions)
 *    M * v (mat4)  = ((M[0]*v.x + M[1]*v.y) + M[2]*v.z) + M[3]*v.w   (columns)
 *    mix(a,b,t)    = a*(1-t) + b*t;    smoothstep = t*t*(3-2t), t = clamp((x-e0)/(e1-e0),0,1)
 *    unorm8 -> f32 = v / 255;  f16 -> f32 exact.
 */
#ifndef HELIO_ORACLE_H
#define HELIO_ORACLE_H

#include <stdint.h>

#include "../include/helio_hip/helio_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host-memory G-buffer: same struct as the library's, data pointers are host. */

/* light_cull.wgsl::main over every tile. */
void or_light_cull_tiles(const GpuCameraUniforms* cam, const LightCullParams* params,
                         const GpuLight* lights, uint32_t* lists, uint32_t* counts);

/* Clustered cull (membership rule of hl_light_cull_clusters).
 * Writes offsets/counts/pool (pool_cap entries) and returns the overflow bits. */
uint32_t or_light_cull_clusters(const GpuCameraUniforms* cam, const hl_cluster_params* cp,
                                const GpuLight* lights, uint32_t num_lights, uint32_t* offsets,
                                uint32_t* counts, uint32_t* pool, uint32_t pool_cap);

/* Raw hit count of one cluster (no cap), for tests. */
uint32_t or_cluster_hits(const GpuCameraUniforms* cam, const hl_cluster_params* cp,
                         const GpuLight* lights, uint32_t num_lights, uint32_t cluster);

/* Same slice rule as the library: largest k < grid_z with depth_split[k] <= depth. */
uint32_t or_slice_of_depth(const hl_cluster_params* cp, float depth);

/* fs_main over the G-buffer (host pointers in gb, host grid arrays in grid).
 * out: RGBA32F, out_pitch bytes.  Sky pixels are left untouched unless
 * HL_SHADE_CLEAR_SKY.  Returns the number of (pixel, light) pairs past the
 * range test (Q).  rows [row0, row1) only (row1 = 0 means all rows). */
uint64_t or_shade(const GpuCameraUniforms* cam, const DeferredGlobals* g, const hl_gbuffer* gb,
                  const GpuLight* lights, uint32_t num_lights, const hl_light_grid* grid,
                  float* out, uint32_t out_pitch, uint32_t flags, uint32_t row0, uint32_t row1);

/* Q for one row range without writing an output. */
uint64_t or_count_pairs(const GpuCameraUniforms* cam, const DeferredGlobals* g,
                        const hl_gbuffer* gb, const GpuLight* lights, uint32_t num_lights,
                        const hl_light_grid* grid);

/* Brute-force conservativeness check of a light grid: for every non-sky pixel
 * and every light that contributes a nonzero term there (FP32 restatement),
 * the light must be in the pixel's cell list, unless that list hit its cap.
 * Returns the number of violating (pixel, light) pairs; *checked gets the
 * number of contributing pairs examined.  Pixels are sampled with `stride`. */
uint64_t or_check_conservative(const GpuCameraUniforms* cam, const DeferredGlobals* g,
                               const hl_gbuffer* gb, const GpuLight* lights, uint32_t num_lights,
                               const hl_light_grid* grid, uint32_t cap, uint32_t stride,
                               uint64_t* checked);

/* World position of pixel (x, y) exactly as fs_main reconstructs it in FP32
 * (deferred_lighting.wgsl:1043-1047). */
void or_world_pos(const GpuCameraUniforms* cam, uint32_t width, uint32_t height, uint32_t x,
                  uint32_t y, float depth, float wp[3]);

/* One pixel-light term of pbr_direct_light in FP32 (exposed for known-answer tests). */
void or_pbr_direct_light(const GpuLight* light, const float world_pos[3], const float N[3],
                         const float V[3], const float F0[3], const float albedo[3],
                         float roughness, float metallic, float out[3]);

/* ------------------------------------------------------------------ FP64 */

/* Decoded per-pixel state the FP64 path works from (one record per pixel). */
typedef struct or64_pixel {
    double wp[3];     /* reconstructed world position (FP64 from the stored depth) */
    double N[3];      /* normalized normal */
    double V[3];      /* normalized view vector */
    double albedo[3];
    double rough, metal, ao;
    double emissive[3];
    double alpha;
    int sky;          /* depth >= 1 */
} or64_pixel;

/* Decode the G-buffer into n = width*height records (FP64). */
void or64_decode(const GpuCameraUniforms* cam, const hl_gbuffer* gb, or64_pixel* px);

/* Forward radiance of one pixel over all lights (F0 resolved from albedo and
 * metallic, as the adjoint differentiates).  active (may be NULL) receives
 * a bitmask word per light chunk -- not used by callers but keeps the FD
 * cut-off detection simple: returns a hash of the active set. */
uint64_t or64_shade_pixel(const or64_pixel* p, const DeferredGlobals* g, const GpuLight* lights,
                          uint32_t num_lights, double out[3]);

/* Full forward: out (n*3 doubles), returns loss against target (RGBA32F host,
 * pitch bytes) if target != NULL. */
double or64_forward(const or64_pixel* px, uint32_t width, uint32_t height,
                    const DeferredGlobals* g, const GpuLight* lights, uint32_t num_lights,
                    double* out, const float* target, uint32_t target_pitch);

/* Analytic adjoint of the L2 loss against target (dL/dout = 2(out - target)):
 * d_pixel: 5 planes of width*height doubles; d_pos: 3 per light; d_int: 1 per light. */
void or64_backward_l2(const or64_pixel* px, uint32_t width, uint32_t height,
                      const DeferredGlobals* g, const GpuLight* lights, uint32_t num_lights,
                      const float* target, uint32_t target_pitch, double* d_pixel,
                      double* d_pos, double* d_int);

/* Same adjoint with a given dL/dout (RGBA32F host, pitch bytes). */
void or64_backward(const or64_pixel* px, uint32_t width, uint32_t height,
                   const DeferredGlobals* g, const GpuLight* lights, uint32_t num_lights,
                   const float* dout, uint32_t dout_pitch, double* d_pixel, double* d_pos,
                   double* d_int);

/* Central finite difference of the L2 loss w.r.t. one parameter.
 * kind 0..4: pixel param (albedo r,g,b, roughness, metallic) of pixel `index`;
 * kind 5..7: light position x,y,z of light `index`; kind 8: light intensity.
 * Returns the FD value; *crosses = 1 when the stencil changes the set of active
 * (pixel, light) terms (a cut-off lies inside [-h, +h]). */
double or64_fd(const or64_pixel* px, uint32_t width, uint32_t height, const DeferredGlobals* g,
               const GpuLight* lights, uint32_t num_lights, const float* target,
               uint32_t target_pitch, int kind, uint32_t index, double h, int* crosses);

/* G-buffer producer contract (gbuffer_pack.c): hl_gbuffer_pack on host
 * memory, and its pieces. */
void or_gbuffer_pack(const hl_surface* surface, const hl_gbuffer* gbuffer);
uint16_t or_float_to_half(float f);
uint32_t or_unorm8_encode(float v);
float or_f0_metallic(float albedo, float metallic);

/* Forward-lit fragment shading (forward_lit.c, forward_lit.wgsl:337-493):
 * hl_forward_lit on host memory (out may be NULL: count only; returns the
 * (fragment, light) pairs past the range test), and its pieces. */
uint64_t or_forward_lit(const GpuCameraUniforms* cam, const ForwardLitGlobals* g, const hl_fragment* frags,
                        uint32_t n, const GpuLight* lights, uint32_t num_lights, const uint32_t* entity,
                        const hl_transform* transforms, uint32_t num_transforms, const hl_light_grid* grid,
                        float* out);
void or_light_direction_from_rotation(const float rot_deg[3], float out[3]);

/* Shadows (shadow.c, deferred_lighting.wgsl:234-565): or_shade with the terms
 * of shadow-casting lights multiplied by shadow_factor (host pointers in sh),
 * and the pieces of shadow_factor. */
uint64_t or_shade_shadowed(const GpuCameraUniforms* cam, const DeferredGlobals* g, const hl_gbuffer* gb,
                           const GpuLight* lights, uint32_t num_lights, const hl_light_grid* grid,
                           const hl_shadow_maps* sh, float* out, uint32_t out_pitch, uint32_t flags,
                           uint32_t row0, uint32_t row1);
/* or_shade with shadows (sh, may be NULL) and the radiance-cascade indirect
 * diffuse from the rc_cascade0 probe atlas (rc, host texture; used when
 * g->has_rc_gi != 0). */
uint64_t or_shade_ex(const GpuCameraUniforms* cam, const DeferredGlobals* g, const hl_gbuffer* gb,
                     const GpuLight* lights, uint32_t num_lights, const hl_light_grid* grid,
                     const hl_shadow_maps* sh, const hl_texture* rc, float* out, uint32_t out_pitch,
                     uint32_t flags, uint32_t row0, uint32_t row1);
/* RC GI pieces (rc_gi.c, deferred_lighting.wgsl:827-904). */
void or_rc_oct_decode(float u, float v, float out[3]);
void or_rc_irradiance(const hl_texture* t, const float wmin[3], const float wmax[3], const float wp[3],
                      const float N[3], float out[3]);

/* reflection.c: fs_reflection (deferred_lighting.wgsl:1266-1341) and its parts */
size_t or_cube_bytes(uint32_t face_size, uint32_t mip_levels, uint32_t format);
void or_cube_face(const float d[3], uint32_t* face, float* u, float* v);
void or_cube_sample(const hl_cube_array* env, uint32_t cube, const float dir[3], float lod, float out[3]);
float or_capture_weight(const float P[3], const GpuReflectionCapture* cap);
void or_capture_parallax(const float P[3], const float R[3], const GpuReflectionCapture* cap, float out[3]);
void or_reflection_env(const DeferredGlobals* g, const hl_reflection_inputs* in, const float P[3], const float R[3],
                       float lod, float out[3]);
void or_env_brdf_approx(float ndv, float roughness, float out[2]);
uint64_t or_reflection(const GpuCameraUniforms* cam, const DeferredGlobals* g, const hl_gbuffer* gb,
                       const hl_reflection_inputs* in, float* out, uint32_t out_pitch, uint32_t flags,
                       uint32_t row0, uint32_t row1);
float or_shadow_factor(const hl_shadow_maps* sh, const GpuLight* lights, uint32_t light_idx, const float cam_pos[3],
                       const float csm_splits[4], const float wp[3], const float N[3], float fx, float fy);
float or_sample_cascade_shadow(const hl_shadow_maps* sh, uint32_t layer, uint32_t cascade, float scale,
                               const float wp[3], float fx, float fy);
float or_sample_cascade_shadow_pcss(const hl_shadow_maps* sh, uint32_t layer, uint32_t cascade, const float wp[3],
                                    float fx, float fy);
float or_sample_compare(const float* atlas, uint32_t size, uint32_t layer, float u, float v, float ref);
void or_sincos(float x, float* s, float* c);
float or_hash22(float px, float py);
void or_vogel(uint32_t i, uint32_t n, float theta, float out[2]);
uint32_t or_point_light_face(const float d[3]);
void or_forward_pbr_direct_light(const GpuLight* light, const float light_pos[3], const float light_dir[3],
                                 const float wp[3], const float N[3], const float V[3], const float F0[3],
                                 const float albedo[3], float roughness, float metallic, float out[3]);

#ifdef __cplusplus
}
#endif
#endif
