// Tests of the C++ pass mirror (include/helio_hip/passes.hpp) through the C ABI.
//
//   test_passes cpu   host-only checks: error behaviour without a GPU, cluster
//                     parameters, the mirror's constants
//   test_passes gpu   a 320x180 scene (reference G-buffer formats, 48 lights,
//                     a quarter of them spots): LightCull in tile and cluster
//                     mode and DeferredLight, run through the pass objects as a
//                     render graph would, against the oracle: lists bit-exact,
//                     radiance within 1e-4 relative + 1e-6 absolute
//
// Prints one line per check and returns 0 when all pass.  The oracle is the
// checker only (test infrastructure).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../helio_amd/csrc/synth.h"
#include "../../oracle/oracle.h"
#include "helio_hip/passes.hpp"

using namespace helio_hip;

static int g_fail = 0;

static void report(bool ok, const std::string& what) {
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", what.c_str());
    if (!ok) ++g_fail;
}

// ------------------------------------------------------------------ cpu
static void test_cpu() {
    // no device here: creating a context fails cleanly with a status, no abort
    try {
        Context ctx(0);
        report(false, "context creation without a GPU throws");
    } catch (const Error& e) {
        report(e.status() == HL_EHIP || e.status() == HL_EINVAL, "context creation without a GPU throws Error");
    }
    try {
        check(nullptr, HL_EINVAL, "probe");
        report(false, "check() throws on a failing status");
    } catch (const Error& e) {
        report(e.status() == HL_EINVAL && std::string(e.what()).find("probe") != std::string::npos,
               "check() throws Error carrying the status and the call");
    }
    // cluster parameters from a host camera
    hl_synth_config c{};
    c.width = 320, c.height = 180, c.fov_y_deg = 60, c.near_plane = 0.1f, c.far_plane = 1000;
    c.eye[1] = 8, c.eye[2] = 32;
    GpuCameraUniforms cam{};
    report(hl_synth_camera(&c, &cam) == 0, "synthetic camera");
    hl_cluster_params p{};
    report(hl_cluster_params_init(&cam, 320, 180, 8, 5, 16, 0, &p) == HL_OK && p.cell_w == 40 && p.cell_h == 36 &&
               p.depth_split[0] == 0.0f && p.depth_split[16] == 1.0f,
           "hl_cluster_params_init cell sizes and depth split ends");
    bool asc = true;
    for (int k = 0; k < 16; ++k) asc = asc && p.depth_split[k] < p.depth_split[k + 1];
    report(asc, "depth splits ascending");
    report(HL_TILE_SIZE == 16u && HL_MAX_LIGHTS_PER_TILE == 64u, "reference tile constants");
}

// ------------------------------------------------------------------ gpu
struct HostScene {
    hl_synth_config cfg{};
    GpuCameraUniforms cam{};
    std::vector<GpuLight> lights;
    std::vector<uint8_t> albedo, normal, orm, emissive;
    std::vector<float> depth;
    DeferredGlobals globals{};
};

static HostScene make_scene() {
    HostScene s;
    hl_synth_config& c = s.cfg;
    c.width = 320, c.height = 180, c.gbuf_format = 0, c.num_lights = 48;
    c.spot_fraction = 0.25f, c.spot_outer_min_deg = 10, c.spot_outer_max_deg = 45, c.spot_inner_ratio = 0.8f;
    const float box_c[3] = {0, 5, 0}, box_s[3] = {40, 10, 40};
    std::memcpy(c.light_box_center, box_c, sizeof box_c);
    std::memcpy(c.light_box_size, box_s, sizeof box_s);
    c.range_min = 2, c.range_max = 10, c.intensity_min = 0.5f, c.intensity_max = 5;
    c.color_min = 0.5f, c.color_max = 1;
    c.eye[0] = 0, c.eye[1] = 8, c.eye[2] = 32;
    c.fov_y_deg = 60, c.near_plane = 0.1f, c.far_plane = 1000;
    c.num_boxes = 40, c.box_field_half = 20, c.box_size_min = 1, c.box_size_max = 4;
    c.box_height_min = 0.5f, c.box_height_max = 6, c.depth_noise = 0.05f;
    c.seed = 0x48454C494F000000ull + 0x777;
    const size_t px = (size_t)c.width * c.height;
    s.lights.resize(c.num_lights);
    s.albedo.resize(4 * px), s.orm.resize(4 * px), s.normal.resize(8 * px), s.emissive.resize(8 * px);
    s.depth.resize(px);
    if (hl_synth_camera(&c, &s.cam) || hl_synth_lights(&c, s.lights.data()) ||
        hl_synth_gbuffer(&c, s.albedo.data(), s.normal.data(), s.orm.data(), s.emissive.data(), s.depth.data()))
        throw Error(HL_EINVAL, "synthetic scene");
    DeferredGlobals& g = s.globals;
    g.light_count = c.num_lights;
    g.ambient_intensity = 1.0f;
    g.ambient_color[0] = 0.5f, g.ambient_color[1] = 0.5f, g.ambient_color[2] = 0.6f, g.ambient_color[3] = 1.0f;
    g.num_tiles_x = (c.width + 15) / 16;
    g.enable_reflections = 1;
    return s;
}

static hl_gbuffer gbuffer_of(uint32_t w, uint32_t h, const void* alb, const void* nrm, const void* orm,
                             const void* emi, const void* dep) {
    hl_gbuffer g{};
    g.width = w, g.height = h;
    g.albedo = {alb, 4 * w, HL_FMT_RGBA8_UNORM};
    g.orm = {orm, 4 * w, HL_FMT_RGBA8_UNORM};
    g.normal = {nrm, 8 * w, HL_FMT_RGBA16_FLOAT};
    g.emissive = {emi, 8 * w, HL_FMT_RGBA16_FLOAT};
    g.depth = {dep, 4 * w, HL_FMT_R32_FLOAT};
    return g;
}

static bool radiance_close(const std::vector<float>& got, const std::vector<float>& ref, const std::vector<float>& depth,
                           double* worst) {
    bool ok = true;
    *worst = 0.0;
    for (size_t p = 0; p < depth.size(); ++p) {
        if (!(depth[p] < 1.0f)) continue;
        for (int k = 0; k < 4; ++k) {
            const double a = got[4 * p + k], b = ref[4 * p + k];
            const double err = std::fabs(a - b), tol = 1e-4 * std::fabs(b) + 1e-6;
            if (err > tol) ok = false;
            if (std::fabs(b) > 1e-3) *worst = std::max(*worst, err / std::fabs(b));
        }
    }
    return ok;
}

static void test_gpu() {
    HostScene hs = make_scene();
    const uint32_t W = hs.cfg.width, H = hs.cfg.height, n = hs.cfg.num_lights;
    Context ctx(0);
    DeviceBuffer cam_d(ctx, sizeof(GpuCameraUniforms)), lights_d(ctx, sizeof(GpuLight) * n);
    DeviceBuffer alb(ctx, hs.albedo.size()), nrm(ctx, hs.normal.size()), orm(ctx, hs.orm.size()),
        emi(ctx, hs.emissive.size()), dep(ctx, 4 * hs.depth.size());
    cam_d.upload(&hs.cam, sizeof hs.cam);
    lights_d.upload(hs.lights.data(), sizeof(GpuLight) * n);
    alb.upload(hs.albedo.data(), hs.albedo.size());
    nrm.upload(hs.normal.data(), hs.normal.size());
    orm.upload(hs.orm.data(), hs.orm.size());
    emi.upload(hs.emissive.data(), hs.emissive.size());
    dep.upload(hs.depth.data(), 4 * hs.depth.size());
    const hl_gbuffer gb = gbuffer_of(W, H, alb.data(), nrm.data(), orm.data(), emi.data(), dep.data());
    const hl_gbuffer gb_host = gbuffer_of(W, H, hs.albedo.data(), hs.normal.data(), hs.orm.data(),
                                          hs.emissive.data(), hs.depth.data());

    SceneView scene;
    scene.camera_d = cam_d.data<GpuCameraUniforms>();
    scene.camera = hs.cam;
    scene.lights_d = lights_d.data<GpuLight>();
    scene.num_lights = n;
    scene.globals = hs.globals;
    PrepareContext prep{&scene, W, H};

    // ---- tile mode, as the default graph wires LightCull -> DeferredLight
    {
        LightCullPass cull(ctx, W, H);
        DeferredLightPass shade(ctx, W, H);
        report(std::string(cull.name()) == "LightCull" && std::string(shade.name()) == "DeferredLight",
               "pass names match the reference");
        const auto w = cull.writes();
        report(w.size() == 2 && std::string(w[0]) == "tile_light_lists" && std::string(w[1]) == "tile_light_counts",
               "LightCull writes the reference slots");
        FrameResources frame;
        frame.gbuffer = &gb;
        PassContext pc{&scene, &frame};
        cull.prepare(prep);
        shade.prepare(prep);
        cull.execute(pc);
        shade.execute(pc);
        ctx.synchronize();
        const LightCullParams& lp = cull.params();
        const size_t tiles = (size_t)lp.num_tiles_x * lp.num_tiles_y;
        std::vector<uint32_t> lists(tiles * HL_MAX_LIGHTS_PER_TILE), counts(tiles);
        cull.lists().download(lists.data(), 4 * lists.size());
        cull.counts().download(counts.data(), 4 * counts.size());
        std::vector<uint32_t> rl(lists.size(), 0), rc(tiles, 0);
        or_light_cull_tiles(&hs.cam, &lp, hs.lights.data(), rl.data(), rc.data());
        bool same = counts == rc;
        for (size_t t = 0; same && t < tiles; ++t)
            for (uint32_t i = 0; i < rc[t]; ++i) same = same && lists[t * 64 + i] == rl[t * 64 + i];
        report(same, "tile lists bit-exact vs oracle");
        std::vector<float> out(4 * (size_t)W * H), ref(4 * (size_t)W * H, 0.0f);
        shade.output().download(out.data(), 4 * out.size());
        hl_light_grid grid{};
        grid.mode = HL_GRID_TILES;
        grid.num_tiles_x = lp.num_tiles_x;
        grid.num_tiles_y = lp.num_tiles_y;
        grid.lists_d = rl.data();
        grid.counts_d = rc.data();
        grid.lists_capacity = (uint32_t)rl.size();
        or_shade(&hs.cam, &shade.globals(), &gb_host, hs.lights.data(), n, &grid, ref.data(), 16 * W, 0, 0, 0);
        double worst = 0;
        const bool ok = radiance_close(out, ref, hs.depth, &worst);
        char buf[96];
        std::snprintf(buf, sizeof buf, "tile-mode radiance within 1e-4 rel (worst %.2e)", worst);
        report(ok, buf);
    }
    // ---- cluster mode
    {
        hl_cluster_params cp{};
        check(nullptr, hl_cluster_params_init(&hs.cam, W, H, 8, 5, 16, 0, &cp), "hl_cluster_params_init");
        const uint32_t pool_cap = 1u << 18, ncell = 8 * 5 * 16;
        LightCullPass cull(ctx, W, H, cp, pool_cap);
        DeferredLightPass shade(ctx, W, H);
        FrameResources frame;
        frame.gbuffer = &gb;
        PassContext pc{&scene, &frame};
        cull.prepare(prep);
        shade.prepare(prep);
        cull.execute(pc);
        shade.execute(pc);
        ctx.synchronize();
        report(cull.overflow() == 0u, "cluster pool large enough");
        std::vector<uint32_t> off(ncell), cnt(ncell), pool(pool_cap);
        cull.offsets().download(off.data(), 4 * off.size());
        cull.counts().download(cnt.data(), 4 * cnt.size());
        cull.lists().download(pool.data(), 4 * pool.size());
        std::vector<uint32_t> ro(ncell), rcnt(ncell), rpool(pool_cap);
        or_light_cull_clusters(&hs.cam, &cp, hs.lights.data(), n, ro.data(), rcnt.data(), rpool.data(), pool_cap);
        bool same = off == ro && cnt == rcnt;
        for (uint32_t c = 0; same && c < ncell; ++c)
            for (uint32_t i = 0; i < rcnt[c]; ++i) same = same && pool[ro[c] + i] == rpool[ro[c] + i];
        report(same, "cluster lists bit-exact vs oracle");
        std::vector<float> out(4 * (size_t)W * H), ref(4 * (size_t)W * H, 0.0f);
        shade.output().download(out.data(), 4 * out.size());
        hl_light_grid grid{};
        grid.mode = HL_GRID_CLUSTERS;
        grid.lists_d = rpool.data();
        grid.counts_d = rcnt.data();
        grid.offsets_d = ro.data();
        grid.clusters = &cp;
        grid.lists_capacity = pool_cap;
        or_shade(&hs.cam, &shade.globals(), &gb_host, hs.lights.data(), n, &grid, ref.data(), 16 * W, 0, 0, 0);
        double worst = 0;
        const bool ok = radiance_close(out, ref, hs.depth, &worst);
        char buf[96];
        std::snprintf(buf, sizeof buf, "cluster-mode radiance within 1e-4 rel (worst %.2e)", worst);
        report(ok, buf);
    }
    // ---- errors surface as exceptions with the library's message
    try {
        DeferredLightPass shade(ctx, W, H);
        FrameResources frame;  // no light grid published
        frame.gbuffer = &gb;
        PassContext pc{&scene, &frame};
        shade.prepare(prep);
        shade.execute(pc);
        report(false, "execute without a light grid throws");
    } catch (const Error& e) {
        report(e.status() == HL_EINVAL, "execute without a light grid throws HL_EINVAL");
    }
    try {
        LightCullPass cull(ctx, W, H);
        PrepareContext bad{&scene, W + 1, H};
        cull.prepare(bad);
        report(false, "prepare with a wrong frame size throws");
    } catch (const Error& e) {
        report(e.status() == HL_EINVAL, "prepare with a wrong frame size throws HL_EINVAL");
    }
}

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    try {
        if (mode == "cpu") test_cpu();
        else if (mode == "gpu") test_gpu();
        else {
            std::fprintf(stderr, "usage: %s cpu|gpu\n", argv[0]);
            return 2;
        }
    } catch (const std::exception& e) {
        std::printf("FAIL unexpected exception: %s\n", e.what());
        return 1;
    }
    std::printf("%s: %d failure(s)\n", mode.c_str(), g_fail);
    return g_fail ? 1 : 0;
}
