// `gargantua --still out.png [--sequence N]` (Still.swift:9-99): settle the
// disk, splat it exactly, march the still with the still policy and write a
// PNG; optionally write N interactive frames with the disk advancing.
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gargantua_constants.h"
#include "host.hpp"

namespace host {

namespace {

std::string sibling(const std::string& path, const std::string& name) {
    size_t slash = path.find_last_of('/');
    return slash == std::string::npos ? name : path.substr(0, slash + 1) + name;
}

void march(garg_ctx* ctx, garg_march_uniforms& u, Scene& s, garg_targets* t, unsigned form) {
    garg_march_request r{};
    r.uniforms = u;
    r.targets = t;
    r.tables = u.spin > 0.0f ? nullptr : s.tables;
    r.volume = s.volume;
    garg_batch_config cfg;
    check(garg_batch_config_default(ctx, &cfg), "reading the batch configuration");
    cfg.kerr_schild = form;
    check(garg_march_batch(ctx, nullptr, &r, 1, &cfg, nullptr), "marching");
}

}  // namespace

void run_still(const Options& o) {
    Configuration c = Configuration::make(o);
    const int width = o.width, height = o.height;
    uint64_t extra = (uint64_t)width * height * 28u;
    Scene s = make_scene(c, (int)STILL_SETTLE_STEPS, true);
    garg_device_info info;
    check(garg_device_info_get(s.ctx, &info), "reading the device");
    std::string refusal = budget_refusal(c.bytes() + extra, info.budget_bytes,
                                         "A " + std::to_string(width) + " by " + std::to_string(height) +
                                             " still at preset " + preset_name(c.preset));
    if (!refusal.empty()) operational(refusal);
    line("settled " + std::to_string(c.particles) + " particles for " + std::to_string(STILL_SETTLE_STEPS) +
         " substeps");
    OrbitCamera camera = camera_from(o);
    garg_march_uniforms u = make_uniforms(camera, width, height, Policy::still, true, (uint32_t)c.seed, true, c.spin,
                                          true);
    if (c.spin == 0.0) {
        check(garg_tables_update(s.tables, nullptr, camera.distance), "updating the sweep tables");
        check(garg_tables_apply(s.tables, &u), "applying the sweep tables");
    }
    garg_targets* full = nullptr;
    check(garg_targets_create(s.ctx, (uint32_t)width, (uint32_t)height, &full), "allocating the still");
    line("rendering " + std::to_string(width) + " by " + std::to_string(height));
    march(s.ctx, u, s, full, o.form);
    std::vector<uint8_t> rgba((size_t)width * height * 4);
    if (o.sequence) {
        const int sw = 960, sh = 540;
        garg_targets* seq = nullptr;
        check(garg_targets_create(s.ctx, sw, sh, &seq), "allocating sequence targets");
        std::vector<uint8_t> frame_rgba((size_t)sw * sh * 4);
        for (int f = 0; f < *o.sequence; ++f) {
            check(garg_particles_step(s.particles, nullptr, 1), "stepping");
            check(garg_splat_bin(s.volume, s.particles, nullptr, 0u, (uint32_t)f), "splatting");
            check(garg_splat_resolve(s.volume, nullptr, 1u), "resolving");
            garg_march_uniforms fu = make_uniforms(camera, sw, sh, Policy::interactive, true, (uint32_t)c.seed, true,
                                                   c.spin, false);
            if (c.spin == 0.0) check(garg_tables_apply(s.tables, &fu), "applying the sweep tables");
            march(s.ctx, fu, s, seq, o.form);
            check(garg_tonemap(seq, nullptr, GARG_F(EXPOSURE_DEFAULT), 0u, 1.0f, frame_rgba.data()), "tone mapping");
            char name[32];
            snprintf(name, sizeof name, "seq-%03d.png", f);
            std::string path = sibling(o.still_path, name);
            if (!write_png(path, frame_rgba.data(), sw, sh)) operational("Could not write the sequence frame to " + path + ".");
        }
        garg_targets_destroy(seq);
        line("sequence of " + std::to_string(*o.sequence) + " frames beside the still");
    }
    check(garg_tonemap(full, nullptr, GARG_F(EXPOSURE_DEFAULT), 0u, 1.0f, rgba.data()), "tone mapping");
    if (!write_png(o.still_path, rgba.data(), width, height)) operational("Could not write the still to " + o.still_path + ".");
    garg_targets_destroy(full);
    s.close();
    line("still " + o.still_path);
    exit(0);
}

}  // namespace host
