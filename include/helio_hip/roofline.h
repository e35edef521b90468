/*
 * roofline.h -- the algorithmic work of one deferred-shading launch, the
 * figures a roofline of the shading kernel is computed from (bench.py reads
 * them from this file).  Per launch over P_sky sky pixels (depth >= 1),
 * P_shaded shaded pixels, N_cells clusters or tiles, E list entries, L lights
 * and Q (pixel, light) pairs past the range test:
 *
 *   bytes = SKY_PIXEL * P_sky
 *         + (G-buffer hot texel bytes + DEPTH + OUTPUT) * P_shaded
 *         + CELL * N_cells + LIST_ENTRY * E + LIGHT * L
 *   flops = FLOP_PER_PIXEL * P_shaded + FLOP_PER_PAIR * Q
 *
 * The G-buffer hot texel bytes are albedo + normal + orm + emissive at the
 * launch's formats (the reference formats 4 + 8 + 4 + 8, fp16 8 each, fp32 16
 * each).  FLOP_PER_PAIR counts the range test (9) and the pbr_direct_light
 * body (121, deferred_lighting.wgsl:699-820 with pbr_eval.wgsl:6-66), each
 * add/mul/cmp/min/max/div/sqrt as 1 and an FMA as 2; FLOP_PER_PIXEL the
 * G-buffer decode, world position, N and V and the ambient combine.
 */
#ifndef HELIO_HIP_ROOFLINE_H
#define HELIO_HIP_ROOFLINE_H

#define HL_RL_SKY_PIXEL_BYTES 4   /* depth of a sky pixel */
#define HL_RL_DEPTH_BYTES 4       /* depth of a shaded pixel */
#define HL_RL_OUTPUT_BYTES 16     /* RGBA32F output */
#define HL_RL_CELL_BYTES 4        /* count per cluster (or tile) */
#define HL_RL_LIST_ENTRY_BYTES 4  /* light index */
#define HL_RL_LIGHT_BYTES 56      /* the hot fields of GpuLight */
#define HL_RL_FLOP_PER_PIXEL 60
#define HL_RL_FLOP_PER_PAIR 130

/* MI355X peaks the fractions are taken against (AMD public figures). */
#define HL_PEAK_HBM_GBS 8000.0
#define HL_PEAK_FP32_TFLOPS 157.3      /* packed FP32 */
#define HL_PEAK_FP32_SCALAR_TFLOPS 78.6 /* one FP32 FMA per lane per cycle */

#endif /* HELIO_HIP_ROOFLINE_H */
