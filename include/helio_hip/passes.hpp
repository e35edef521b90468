// helio_hip/passes.hpp -- C++ host-side mirror of Helio's pass interface for the
// deferred-lighting hot path, over the C ABI of helio_hip.h (header-only,
// C++17, no HIP headers needed).
//
// It keeps the shape of the reference's passes so that code driving them reads
// like the Rust renderer:
//   RenderPass          crates/helio-core/src/traits.rs:315-530 (name, reads,
//                       writes, prepare, execute, publish)
//   PrepareContext      crates/helio-core/src/context.rs:503-547 (width, height, scene)
//   PassContext         crates/helio-core/src/context.rs:192-260 (scene, frame resources)
//   FrameResources      crates/libhelio/src/frame.rs:188-480 (the light-grid slots
//                       tile_light_lists / tile_light_counts :230-234 and the G-buffer)
//   LightCullPass       crates/passes/3d/helio-pass-light-cull/src/lib.rs:41-363
//   DeferredLightPass   crates/passes/3d/helio-pass-deferred-light/src/lib.rs:54-1230
// Errors are thrown as helio_hip::Error carrying the hl_status and the
// context's message, as the reference returns Error::InvalidPassConfig
// (crates/helio-core/src/error.rs:62) from prepare/execute.
#ifndef HELIO_HIP_PASSES_HPP
#define HELIO_HIP_PASSES_HPP

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "helio_hip.h"

namespace helio_hip {

class Error : public std::runtime_error {
public:
    Error(hl_status status, const std::string& msg) : std::runtime_error(msg), status_(status) {}
    hl_status status() const { return status_; }

private:
    hl_status status_;
};

inline void check(const hl_context* ctx, hl_status s, const char* what) {
    if (s == HL_OK) return;
    std::string msg = what;
    msg += ": ";
    msg += hl_status_string(s);
    if (ctx) {
        const char* e = hl_last_error(ctx);
        if (e && *e) {
            msg += ": ";
            msg += e;
        }
    }
    throw Error(s, msg);
}

// One device, one stream, the library's scratch (hl_context).
class Context {
public:
    explicit Context(int device = 0) { check(nullptr, hl_context_create(device, &ctx_), "hl_context_create"); }
    ~Context() {
        if (ctx_) (void)hl_context_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    hl_context* get() const { return ctx_; }
    void set_stream(void* stream) { check(ctx_, hl_set_stream(ctx_, stream), "hl_set_stream"); }
    void synchronize() { check(ctx_, hl_synchronize(ctx_), "hl_synchronize"); }

private:
    hl_context* ctx_ = nullptr;
};

// A device allocation (hl_alloc / hl_free).
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(Context& ctx, size_t bytes) : ctx_(&ctx), bytes_(bytes) {
        check(ctx.get(), hl_alloc(ctx.get(), bytes ? bytes : 16, &ptr_), "hl_alloc");
    }
    ~DeviceBuffer() { reset(); }
    DeviceBuffer(DeviceBuffer&& o) noexcept { *this = std::move(o); }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            reset();
            std::swap(ctx_, o.ctx_);
            std::swap(ptr_, o.ptr_);
            std::swap(bytes_, o.bytes_);
        }
        return *this;
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;

    template <class T = void>
    T* data() const { return static_cast<T*>(ptr_); }
    size_t bytes() const { return bytes_; }
    void upload(const void* src, size_t n) {
        if (n > bytes_) throw Error(HL_EINVAL, "upload larger than the buffer");
        check(ctx_->get(), hl_memcpy_htod(ctx_->get(), ptr_, src, n), "hl_memcpy_htod");
    }
    void download(void* dst, size_t n) const {
        if (n > bytes_) throw Error(HL_EINVAL, "download larger than the buffer");
        check(ctx_->get(), hl_memcpy_dtoh(ctx_->get(), dst, ptr_, n), "hl_memcpy_dtoh");
    }
    void fill_u32(uint32_t value) {
        check(ctx_->get(), hl_memset_d32(ctx_->get(), ptr_, value, bytes_ / 4), "hl_memset_d32");
    }

private:
    void reset() {
        if (ptr_ && ctx_) (void)hl_free(ctx_->get(), ptr_);
        ptr_ = nullptr;
    }
    Context* ctx_ = nullptr;
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

// What the passes read from the scene (SceneResources): device camera and
// lights, a host copy of the camera, and the frame's DeferredGlobals.
struct SceneView {
    const GpuCameraUniforms* camera_d = nullptr;
    GpuCameraUniforms camera{};
    const GpuLight* lights_d = nullptr;
    uint32_t num_lights = 0;
    DeferredGlobals globals{};
};

struct PrepareContext {
    const SceneView* scene = nullptr;
    uint32_t width = 0, height = 0;
};

// The frame slots the two passes exchange (FrameResources): the light grid
// published by LightCull, the G-buffer read by DeferredLight and its pre_aa
// target.
struct FrameResources {
    hl_light_grid light_grid{};
    bool has_light_grid = false;
    const hl_gbuffer* gbuffer = nullptr;
    float* pre_aa_d = nullptr;
    uint32_t pre_aa_pitch = 0;
};

struct PassContext {
    const SceneView* scene = nullptr;
    FrameResources* frame = nullptr;
};

class RenderPass {
public:
    virtual ~RenderPass() = default;
    virtual const char* name() const = 0;
    virtual std::vector<const char*> reads() const { return {}; }
    virtual std::vector<const char*> writes() const { return {}; }
    virtual void prepare(const PrepareContext&) {}
    virtual void execute(PassContext& ctx) = 0;
    virtual void publish(FrameResources&) const {}
};

// LightCull: the reference's 16-px tile lists (same lists, order and cap), or
// log-depth clustered lists when constructed with cluster parameters.
class LightCullPass : public RenderPass {
public:
    // tile mode (reference layout: lists[tile * 64 + i], counts[tile])
    LightCullPass(Context& ctx, uint32_t width, uint32_t height)
        : ctx_(ctx), width_(width), height_(height), tiles_x_((width + HL_TILE_SIZE - 1) / HL_TILE_SIZE),
          tiles_y_((height + HL_TILE_SIZE - 1) / HL_TILE_SIZE),
          lists_(ctx, 4 * (size_t)tiles_x_ * tiles_y_ * HL_MAX_LIGHTS_PER_TILE),
          counts_(ctx, 4 * (size_t)tiles_x_ * tiles_y_) {}
    // cluster mode (compact lists: offsets / counts / pool of pool_cap entries)
    LightCullPass(Context& ctx, uint32_t width, uint32_t height, const hl_cluster_params& params,
                  uint32_t pool_cap)
        : ctx_(ctx), width_(width), height_(height), clustered_(true), cparams_(params), pool_cap_(pool_cap),
          lists_(ctx, 4 * (size_t)(pool_cap ? pool_cap : 1)),
          counts_(ctx, 4 * (size_t)params.grid_x * params.grid_y * params.grid_z),
          offsets_(ctx, 4 * (size_t)params.grid_x * params.grid_y * params.grid_z), overflow_(ctx, 16) {}

    const char* name() const override { return "LightCull"; }
    std::vector<const char*> writes() const override {
        if (clustered_) return {"cluster_light_grid"};
        return {"tile_light_lists", "tile_light_counts"};
    }

    // LightCullParams as LightCullPass::prepare fills them (lib.rs:264-282)
    void prepare(const PrepareContext& pc) override {
        if (!pc.scene) throw Error(HL_EINVAL, "LightCull::prepare: no scene");
        if (pc.width != width_ || pc.height != height_)
            throw Error(HL_EINVAL, "LightCull::prepare: frame size differs from the pass's");
        std::memset(&params_, 0, sizeof params_);
        params_.num_tiles_x = (width_ + HL_TILE_SIZE - 1) / HL_TILE_SIZE;
        params_.num_tiles_y = (height_ + HL_TILE_SIZE - 1) / HL_TILE_SIZE;
        params_.num_lights = pc.scene->num_lights;
        params_.screen_width = width_;
        params_.screen_height = height_;
    }

    void execute(PassContext& pc) override {
        const SceneView& s = *pc.scene;
        if (!clustered_) {
            check(ctx_.get(),
                  hl_light_cull_tiles(ctx_.get(), s.camera_d, &params_, s.lights_d, lists_.data<uint32_t>(),
                                      counts_.data<uint32_t>()),
                  "hl_light_cull_tiles");
        } else {
            check(ctx_.get(),
                  hl_light_cull_clusters(ctx_.get(), s.camera_d, &cparams_, s.lights_d, s.num_lights,
                                         offsets_.data<uint32_t>(), counts_.data<uint32_t>(),
                                         lists_.data<uint32_t>(), pool_cap_, overflow_.data<uint32_t>()),
                  "hl_light_cull_clusters");
        }
        if (pc.frame) publish(*pc.frame);
    }

    void publish(FrameResources& f) const override {
        hl_light_grid g{};
        if (!clustered_) {
            g.mode = HL_GRID_TILES;
            g.num_tiles_x = tiles_x_;
            g.num_tiles_y = tiles_y_;
            g.lists_d = lists_.data<uint32_t>();
            g.counts_d = counts_.data<uint32_t>();
            g.lists_capacity = tiles_x_ * tiles_y_ * HL_MAX_LIGHTS_PER_TILE;
        } else {
            g.mode = HL_GRID_CLUSTERS;
            g.lists_d = lists_.data<uint32_t>();
            g.counts_d = counts_.data<uint32_t>();
            g.offsets_d = offsets_.data<uint32_t>();
            g.clusters = &cparams_;
            g.lists_capacity = pool_cap_;
        }
        f.light_grid = g;
        f.has_light_grid = true;
    }

    // HL_OVERFLOW_* bits of the last clustered call (synchronises)
    uint32_t overflow() const {
        uint32_t v = 0;
        check(ctx_.get(), hl_query_overflow(ctx_.get(), &v), "hl_query_overflow");
        return v;
    }
    const DeviceBuffer& lists() const { return lists_; }
    const DeviceBuffer& counts() const { return counts_; }
    const DeviceBuffer& offsets() const { return offsets_; }
    const LightCullParams& params() const { return params_; }

private:
    Context& ctx_;
    uint32_t width_, height_;
    bool clustered_ = false;
    uint32_t tiles_x_ = 0, tiles_y_ = 0;
    hl_cluster_params cparams_{};
    uint32_t pool_cap_ = 0;
    LightCullParams params_{};
    DeviceBuffer lists_, counts_, offsets_, overflow_;
};

// DeferredLight: fs_main over the G-buffer into an RGBA32F pre_aa target the
// pass owns; plus the adjoint pair of the differentiable path.
class DeferredLightPass : public RenderPass {
public:
    DeferredLightPass(Context& ctx, uint32_t width, uint32_t height)
        : ctx_(ctx), width_(width), height_(height), out_(ctx, 16 * (size_t)width * height) {}

    const char* name() const override { return "DeferredLight"; }
    std::vector<const char*> reads() const override {
        return {"gbuffer", "tile_light_lists", "tile_light_counts"};
    }
    std::vector<const char*> writes() const override { return {"pre_aa"}; }

    void set_debug_mode(uint32_t mode) { debug_mode_ = mode; }
    void set_clear_sky(bool on) {
        flags_ = on ? (flags_ | (uint32_t)HL_SHADE_CLEAR_SKY) : (flags_ & ~(uint32_t)HL_SHADE_CLEAR_SKY);
    }
    // Use the light records the frame's LightCull (same context, same lights)
    // already built, instead of building them again.
    void set_lights_from_cull(bool on) {
        flags_ = on ? (flags_ | (uint32_t)HL_SHADE_LIGHTS_FROM_CULL)
                    : (flags_ & ~(uint32_t)HL_SHADE_LIGHTS_FROM_CULL);
    }

    // DeferredGlobals as DeferredLightPass::prepare writes them (lib.rs:836-878)
    void prepare(const PrepareContext& pc) override {
        if (!pc.scene) throw Error(HL_EINVAL, "DeferredLight::prepare: no scene");
        if (pc.width != width_ || pc.height != height_)
            throw Error(HL_EINVAL, "DeferredLight::prepare: frame size differs from the pass's");
        globals_ = pc.scene->globals;
        globals_.debug_mode = debug_mode_;
        globals_.num_tiles_x = (width_ + HL_TILE_SIZE - 1) / HL_TILE_SIZE;
        globals_.light_count = pc.scene->num_lights;
    }

    void execute(PassContext& pc) override {
        const FrameResources& f = frame_of(pc);
        check(ctx_.get(),
              hl_deferred_forward(ctx_.get(), pc.scene->camera_d, &globals_, f.gbuffer, pc.scene->lights_d,
                                  pc.scene->num_lights, &f.light_grid, out_.data<float>(), 16 * width_, flags_),
              "hl_deferred_forward");
        publish(*pc.frame);
    }

    void publish(FrameResources& f) const override {
        f.pre_aa_d = out_.data<float>();
        f.pre_aa_pitch = 16 * width_;
    }

    // Forward + per-pixel L2 loss against target_d (RGBA32F) + gradients.
    // d_pixel_d: 5 planes of width*height; d_pos_d: 3 per light; d_int_d: 1 per light.
    void forward_backward_l2(PassContext& pc, const float* target_d, double* loss_d, float* d_pixel_d,
                             float* d_pos_d, float* d_int_d) {
        const FrameResources& f = frame_of(pc);
        check(ctx_.get(),
              hl_deferred_forward_backward_l2(ctx_.get(), pc.scene->camera_d, &globals_, f.gbuffer,
                                              pc.scene->lights_d, pc.scene->num_lights, &f.light_grid, target_d,
                                              16 * width_, out_.data<float>(), 16 * width_, loss_d, d_pixel_d,
                                              d_pos_d, d_int_d),
              "hl_deferred_forward_backward_l2");
    }

    // Gradients for a given dL/dout (RGBA32F, alpha ignored).
    void backward(PassContext& pc, const float* dL_dout_d, float* d_pixel_d, float* d_pos_d, float* d_int_d) {
        const FrameResources& f = frame_of(pc);
        check(ctx_.get(),
              hl_deferred_backward(ctx_.get(), pc.scene->camera_d, &globals_, f.gbuffer, pc.scene->lights_d,
                                   pc.scene->num_lights, &f.light_grid, dL_dout_d, 16 * width_, d_pixel_d,
                                   d_pos_d, d_int_d),
              "hl_deferred_backward");
    }

    const DeviceBuffer& output() const { return out_; }
    const DeferredGlobals& globals() const { return globals_; }

private:
    static const FrameResources& frame_of(const PassContext& pc) {
        if (!pc.scene || !pc.frame) throw Error(HL_EINVAL, "DeferredLight: no scene or frame");
        if (!pc.frame->has_light_grid) throw Error(HL_EINVAL, "DeferredLight: no light grid published");
        if (!pc.frame->gbuffer) throw Error(HL_EINVAL, "DeferredLight: no G-buffer");
        return *pc.frame;
    }
    Context& ctx_;
    uint32_t width_, height_;
    uint32_t debug_mode_ = 0, flags_ = 0;
    DeferredGlobals globals_{};
    DeviceBuffer out_;
};

}  // namespace helio_hip

#endif  // HELIO_HIP_PASSES_HPP
