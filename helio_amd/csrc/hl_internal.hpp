// hl_internal.hpp -- declarations shared by the kernel and API translation units.
#pragma once

#include <hip/hip_runtime.h>

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

// launchers (hl_cull.hip)
hipError_t launch_prep(hipStream_t s, const GpuLight* lights, const uint32_t* subset, uint32_t n,
                       const GpuCameraUniforms* cam, HotLight* hot, CullLight* cull,
                       uint32_t* flags);
hipError_t launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n);
hipError_t launch_cull_tiles(hipStream_t s, const GpuCameraUniforms* cam, const LightCullParams& p,
                             const CullLight* cl, uint32_t* lists, uint32_t* counts);
// Scratch of the clustered cull.  bits (ncell * ceil(L/32) words) and summary
// (ncell * ceil(L/1024) words) must be zero before the first call; every call
// leaves them zero again.
struct ClusterScratch {
    void* geom;          // cluster_geom_bytes(columns, cells)
    uint32_t* bits;
    uint32_t* summary;
    uint32_t* raw;       // ncell
};
size_t cluster_geom_bytes(uint32_t ncol, uint32_t ncell);
size_t cluster_geom_offset_boxes(uint32_t ncol);
inline size_t cluster_bits_words(uint32_t ncell, uint32_t num_lights) {
    return (size_t)ncell * ((num_lights + 31u) / 32u);
}
inline size_t cluster_summary_words(uint32_t ncell, uint32_t num_lights) {
    return (size_t)ncell * (((num_lights + 31u) / 32u + 31u) / 32u);
}
hipError_t launch_cluster_cull(hipStream_t s, const ClusterArgs& a, const GpuCameraUniforms* cam,
                               const CullLight* cl, const ClusterScratch& w, uint32_t* offsets,
                               uint32_t* counts, uint32_t* pool, uint32_t pool_cap,
                               uint32_t* flags);

// launchers (hl_shade.hip)
enum ShadeMode : int { kModeForward = 0, kModeCount = 1, kModeBackward = 2, kModeFusedL2 = 3 };
hipError_t launch_shade(hipStream_t s, int gbuf_set, int mode, const ShadeArgs& a);
hipError_t launch_zero_f64(hipStream_t s, double* p, size_t n);
hipError_t launch_light_grads_to_f32(hipStream_t s, const double* acc, uint32_t n,
                                     const uint32_t* subset, float* d_pos, float* d_int);

}  // namespace hl
