// `gargantua bench [--sweep]` (Bench.swift:126-266): times each pass of the
// default scene with 5 warm-ups and 30 iterations, each iteration bracketed
// by events on the stream it runs on, prints a verdict and writes a JSON
// report under bench/results. --sweep runs the batch autotuner over block
// size, compaction period and refill threshold and persists the winner.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <functional>
#include <map>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/gargantua_constants.h"
#include "host.hpp"

namespace host {

namespace {

struct Timing {
    double mean = 0, median = 0, maximum = 0;
    explicit Timing(std::vector<double> ms = {}) {
        if (ms.empty()) return;
        for (double m : ms) mean += m;
        mean /= (double)ms.size();
        std::sort(ms.begin(), ms.end());
        median = ms[ms.size() / 2];
        maximum = ms.back();
    }
};

const int kWarmup = 5, kIterations = 30;

Timing time_pass(garg_ctx* ctx, const std::function<void()>& encode, int iterations = kIterations) {
    garg_event *a = nullptr, *b = nullptr;
    check(garg_event_create(ctx, &a), "creating an event");
    check(garg_event_create(ctx, &b), "creating an event");
    std::vector<double> samples;
    for (int i = 0; i < kWarmup + iterations; ++i) {
        check(garg_event_record(a, nullptr), "recording");
        encode();
        check(garg_event_record(b, nullptr), "recording");
        check(garg_event_sync(b), "waiting");
        float ms = 0;
        check(garg_event_elapsed_ms(a, b, &ms), "timing");
        if (i >= kWarmup) samples.push_back(ms);
    }
    garg_event_destroy(a);
    garg_event_destroy(b);
    return Timing(samples);
}

std::string json_timing(const Timing& t) {
    char buf[160];
    snprintf(buf, sizeof buf, "{\"max_ms\": %.4f, \"mean_ms\": %.4f, \"p50_ms\": %.4f}", t.maximum, t.mean, t.median);
    return buf;
}

}  // namespace

// The autotuner: every block size x compaction period x refill threshold on
// the 1080p spin 0.9 still frame; the fastest is written to `path`.
static void sweep_batch(Scene& s, const std::string& path) {
    OrbitCamera cam;
    cam.distance = 50.0f;
    cam.elevation = (float)(30.0 * M_PI / 180.0);
    cam.azimuth = 0.0f;
    cam.fov_y = (float)(60.0 * M_PI / 180.0);
    garg_march_uniforms u = make_uniforms(cam, 1920, 1080, Policy::still, true, 1u, true, 0.9, true);
    garg_targets* t = nullptr;
    check(garg_targets_create(s.ctx, 1920, 1080, &t), "allocating targets");
    garg_march_request r{};
    r.uniforms = u;
    r.targets = t;
    r.volume = s.volume;
    garg_batch_config best{};
    double best_ms = 1e30;
    line("batch autotuner: 1920x1080, spin 0.9, 4096-step cap; block, waves/SIMD, period, threshold, ms");
    for (uint32_t block : {64u, 128u, 256u})
        for (uint32_t waves : {2u, 4u, 8u})
            for (uint32_t period : {8u, 16u, 32u, 64u, 128u})
                for (uint32_t threshold : {16u, 32u, 48u, 63u}) {
                    garg_batch_config c{block, waves, period, threshold, 0u};
                    Timing tm = time_pass(
                        s.ctx, [&] { check(garg_march_batch(s.ctx, nullptr, &r, 1, &c, nullptr), "marching"); }, 5);
                    char buf[128];
                    snprintf(buf, sizeof buf, "%3u %u %3u %2u  %8.3f", block, waves, period, threshold, tm.mean);
                    line(buf);
                    if (tm.mean < best_ms) {
                        best_ms = tm.mean;
                        best = c;
                    }
                }
    FILE* f = fopen(path.c_str(), "w");
    if (f) {
        garg_device_info info;
        check(garg_device_info_get(s.ctx, &info), "reading the device");
        fprintf(f,
                "{\"arch\": \"%s\", \"workload\": \"kerr_march_1920x1080_a0.9_4096\", \"block\": %u, "
                "\"waves_per_simd\": %u, \"compaction_period\": %u, \"refill_threshold\": %u, \"ms\": %.4f}\n",
                info.arch, best.block, best.waves_per_simd, best.compaction_period, best.refill_threshold, best_ms);
        fclose(f);
    }
    char buf[160];
    snprintf(buf, sizeof buf, "best: block %u, %u waves/SIMD, period %u, threshold %u: %.3f ms -> %s", best.block,
             best.waves_per_simd, best.compaction_period, best.refill_threshold, best_ms, path.c_str());
    line(buf);
    garg_targets_destroy(t);
}

void run_bench(const Options& o) {
    Configuration c = Configuration::make(o);
    if (c.spin > 0.0) usage_error("The bench measures the Schwarzschild pipeline; run it without spin.");
    c.strategy = Strategy::bake;
    Scene s = make_scene(c, (int)STILL_SETTLE_STEPS, false);
    if (o.sweep) {
        sweep_batch(s, "gargantua_amd/autotune.json");
        s.close();
        exit(0);
    }
    const int w = c.march_width, h = c.march_height;
    OrbitCamera cam;
    garg_march_uniforms u = make_uniforms(cam, w, h, Policy::interactive, true, (uint32_t)c.seed, true, 0.0, false);
    u.bake_spacing = (float)(2.0 * VOLUME_HALF_EXTENT / c.volume * BAKE_SPACING_VOXELS);
    check(garg_tables_update(s.tables, nullptr, cam.distance), "updating tables");
    check(garg_tables_apply(s.tables, &u), "applying tables");
    garg_march_uniforms bake_u = u;
    bake_u.geodesic = policy_settings(Policy::still);
    garg_targets* t = nullptr;
    check(garg_targets_create(s.ctx, w, h, &t), "allocating targets");
    garg_bake* bake = nullptr;
    check(garg_bake_create(s.ctx, w, h, BAKE_SAMPLES, &bake), "allocating bake buffers");
    std::vector<uint8_t> rgba((size_t)w * h * 4);
    line("timing " + std::to_string(kIterations) + " iterations per pass after " + std::to_string(kWarmup) +
         " warmups, " + std::to_string(w) + " by " + std::to_string(h) + " march");
    std::map<std::string, Timing> passes;
    auto report = [&](const std::string& name, const Timing& tm) {
        passes[name] = tm;
        char buf[160];
        snprintf(buf, sizeof buf, "%-14s mean %7.3f ms  p50 %7.3f ms  max %7.3f ms", name.c_str(), tm.mean, tm.median,
                 tm.maximum);
        line(buf);
    };
    uint32_t frame = 0;
    report("sim", time_pass(s.ctx, [&] { check(garg_particles_step(s.particles, nullptr, 1), "stepping"); }));
    report("splat", time_pass(s.ctx, [&] {
               check(garg_splat_bin(s.volume, s.particles, nullptr, 0u, frame++), "splatting");
               check(garg_splat_resolve(s.volume, nullptr, 1u), "resolving");
           }));
    report("march", time_pass(s.ctx, [&] {
               check(garg_march_image(s.ctx, nullptr, &u, s.volume, s.tables, t, 0, 0, GARG_INTEGRATOR_REFERENCE),
                     "marching");
           }));
    garg_march_request r{};
    r.uniforms = u;
    r.targets = t;
    r.tables = s.tables;
    r.volume = s.volume;
    garg_batch_config cfg;
    check(garg_batch_config_default(s.ctx, &cfg), "reading the batch configuration");
    report("march_batched",
           time_pass(s.ctx, [&] { check(garg_march_batch(s.ctx, nullptr, &r, 1, &cfg, nullptr), "marching"); }));
    report("bake", time_pass(s.ctx, [&] {
               int rows = std::max(h / (int)BAKE_FRAMES, 1);
               for (int row = 0; row < h; row += rows)
                   check(garg_bake_rows(bake, nullptr, &bake_u, s.tables, row, std::min(rows, h - row)), "baking");
           }));
    report("walk", time_pass(s.ctx, [&] { check(garg_bake_walk(bake, nullptr, &u, s.volume, t), "walking"); }));
    report("present", time_pass(s.ctx, [&] {
               check(garg_tonemap(t, nullptr, GARG_F(EXPOSURE_DEFAULT), 0u, 1.0f, rgba.data()), "tone mapping");
           }));
    uint32_t overflow = 0;
    check(garg_bake_overflow(bake, &overflow), "reading the overflow count");
    double shared = passes["sim"].mean + passes["splat"].mean + passes["present"].mean;
    double march_frame = shared + passes["march_batched"].mean, bake_frame = shared + passes["walk"].mean;
    uint64_t bake_bytes = garg_bake_bytes(w, h, BAKE_SAMPLES);
    char verdict[512];
    snprintf(verdict, sizeof verdict,
             "march frame %.2f ms, bake frame %.2f ms after a %.2f ms bake per camera stop and %s of samples (%u rays "
             "over the sample budget); %s",
             march_frame, bake_frame, passes["bake"].mean, format_bytes(bake_bytes).c_str(), overflow,
             bake_frame < march_frame ? "the walk is faster while the camera holds still, the march is free to move"
                                      : "the march is faster");
    line(verdict);
    time_t now = time(nullptr);
    char stamp[32];
    strftime(stamp, sizeof stamp, "%Y-%m-%d-%H%M%S", gmtime(&now));
    mkdir("bench", 0755);
    mkdir("bench/results", 0755);
    std::string path = std::string("bench/results/") + stamp + ".json";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) operational("Could not write the bench report to " + path + ".");
    garg_device_info info;
    check(garg_device_info_get(s.ctx, &info), "reading the device");
    fprintf(f, "{\n  \"bake\": {\"bytes\": %llu, \"overflow_rays\": %u},\n  \"date\": \"%s\",\n  \"device\": \"%s %s\",\n",
            (unsigned long long)bake_bytes, overflow, stamp, info.name, info.arch);
    fprintf(f, "  \"frames\": {\"bake_ms\": %.4f, \"march_ms\": %.4f},\n  \"iterations\": %d,\n", bake_frame,
            march_frame, kIterations);
    fprintf(f, "  \"march\": [%d, %d],\n  \"particles\": %d,\n  \"passes\": {\n", w, h, c.particles);
    size_t k = 0;
    for (auto& [name, tm] : passes)
        fprintf(f, "    \"%s\": %s%s\n", name.c_str(), json_timing(tm).c_str(), ++k < passes.size() ? "," : "");
    fprintf(f, "  },\n  \"preset\": \"%s\",\n  \"verdict\": \"%s\",\n  \"version\": \"%s\",\n  \"volume\": %d\n}\n",
            preset_name(c.preset), verdict, garg_version(), c.volume);
    fclose(f);
    line("report " + path);
    garg_bake_destroy(bake);
    garg_targets_destroy(t);
    s.close();
    exit(0);
}

}  // namespace host
