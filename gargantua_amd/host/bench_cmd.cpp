// `gargantua bench [--sweep | --sweep-batch]` (Bench.swift:126-266): times
// each pass of the default scene with 5 warm-ups and 30 iterations, each
// iteration bracketed by events on the stream it runs on, prints a verdict
// and writes a JSON report under bench/results in the reference's schema
// (Bench.swift:236-264). --sweep is the reference's step-policy, splat and
// launch-shape sweep (Bench.swift:49-124); --sweep-batch runs the batch
// autotuner over the persistent march's launch and persists the winner.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <functional>
#include <map>
#include <string>
#include <sys/stat.h>
#include <sys/utsname.h>
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

// Luminance of every pixel of an rgba16f colour image (LensedRender.luminance).
static std::vector<double> luminance(garg_targets* t, int w, int h) {
    std::vector<uint16_t> color((size_t)w * h * 4);
    check(garg_targets_read(t, color.data(), nullptr, nullptr), "reading the image");
    std::vector<double> out((size_t)w * h);
    for (size_t i = 0; i < out.size(); ++i)
        out[i] = 0.2126 * half_to_float(color[4 * i]) + 0.7152 * half_to_float(color[4 * i + 1]) +
                 0.0722 * half_to_float(color[4 * i + 2]);
    return out;
}

// `bench --sweep` (Bench.swift:49-124): the step policies of the campaign,
// each timed on the plain march (reference form) and judged by the drift of
// a 32 x 32 grid of camera probes at the interactive step cap (with the
// sphere jump), the exhausted probes and the image difference from the spec
// policy without stars; then the splat's phases; then the launch shapes.
// The reference sweeps Metal threadgroup shapes; here one 8 x 8 tile is one
// wave64, so the launch-shape sweep is the batched march's workgroup size and
// resident waves per SIMD at the spec policy.
static void sweep(Scene& s, const Configuration& c) {
    const int w = c.march_width, h = c.march_height;
    OrbitCamera cam;
    check(garg_tables_update(s.tables, nullptr, cam.distance), "updating tables");
    const uint32_t cap = STEP_CAP_INTERACTIVE;
    const double voxel = 2.0 * VOLUME_HALF_EXTENT / c.volume;
    struct Named {
        std::string name;
        garg_geodesic_settings g;
    };
    auto policy = [&](const char* name, double factor, double lo, double hi, double empty_factor = -1,
                      double empty_max = -1) {
        garg_geodesic_settings g = policy_settings(Policy::interactive);
        g.step_factor = (float)factor;
        g.step_min = (float)lo;
        g.step_max = (float)hi;
        g.step_cap = cap;
        g.empty_step_factor = (float)(empty_factor < 0 ? factor : empty_factor);
        g.empty_step_max = (float)(empty_max < 0 ? hi : empty_max);
        return Named{name, g};
    };
    const std::vector<Named> policies = {
        policy("fixed 0.25", 1e9, 0.25, 0.25),
        policy("fixed 0.5", 1e9, 0.5, 0.5),
        policy("prop 0.02 max 0.5 (spec)", 0.02, 0.02, 0.5),
        policy("prop 0.03 max 0.5", 0.03, 0.02, 0.5),
        policy("prop 0.04 max 0.5", 0.04, 0.02, 0.5),
        policy("prop 0.05 max 1.0", 0.05, 0.02, 1.0),
        policy("prop 0.02, empty 0.06 max 1.5", 0.02, 0.02, 0.5, 0.06, 1.5),
        policy("prop 0.02, empty 0.1 max 3.0", 0.02, 0.02, 0.5, 0.1, 3.0),
        policy("prop 0.02, empty 0.15 max 4.0", 0.02, 0.02, 0.5, 0.15, 4.0),
        policy("prop 0.02, empty 0.2 max 6.0", 0.02, 0.02, 0.5, 0.2, 6.0),
        policy("prop 0.04, empty 0.08 max 2.0", 0.04, 0.02, 0.5, 0.08, 2.0),
        policy("prop 0.04 max voxel, empty 0.1 max 3", 0.04, 0.02, voxel, 0.1, 3.0),
        policy("prop 0.05 max voxel, empty 0.1 max 3", 0.05, 0.02, voxel, 0.1, 3.0),
        policy("prop 0.05 max 1.0, empty 0.1 max 3", 0.05, 0.02, 1.0, 0.1, 3.0),
        policy("prop 0.06 max 1.0, empty 0.12 max 3", 0.06, 0.02, 1.0, 0.12, 3.0),
    };
    const std::vector<garg_ray_probe> launches = camera_grid(cam, 32, 32, (double)w / h);
    garg_targets* t = nullptr;
    check(garg_targets_create(s.ctx, w, h, &t), "allocating targets");
    auto uniforms_for = [&](const garg_geodesic_settings& g, bool stars) {
        garg_march_uniforms u =
            make_uniforms(cam, w, h, Policy::interactive, true, (uint32_t)c.seed, stars, 0.0, false);
        u.geodesic = g;
        u.drift_budget = GARG_F(DRIFT_BUDGET_INTERACTIVE);
        check(garg_tables_apply(s.tables, &u), "applying tables");
        return u;
    };
    auto render = [&](const garg_geodesic_settings& g) {
        garg_march_uniforms u = uniforms_for(g, false);
        check(garg_march_image(s.ctx, nullptr, &u, s.volume, s.tables, t, 0, 0, GARG_INTEGRATOR_REFERENCE), "marching");
        check(garg_stream_sync(s.ctx, nullptr), "waiting");
        return luminance(t, w, h);
    };
    const std::vector<double> spec = render(policies[2].g);
    line("step policy sweep: 15 march iterations each, drift over " + std::to_string(launches.size()) +
         " camera rays at the " + std::to_string(cap) +
         " step cap, image difference from the spec policy without stars");
    char buf[256];
    snprintf(buf, sizeof buf, "%-38s  march ms  max drift  p99 drift  exhausted  image diff", "policy");
    line(buf);
    for (const Named& p : policies) {
        garg_march_uniforms u = uniforms_for(p.g, true);
        Timing tm = time_pass(
            s.ctx,
            [&] { check(garg_march_image(s.ctx, nullptr, &u, s.volume, s.tables, t, 0, 0, GARG_INTEGRATOR_REFERENCE),
                        "marching"); },
            15);
        garg_march_uniforms pu = u;
        pu.drift_budget = 1.0f;
        pu.redshift = 0u;
        std::vector<garg_ray_probe_result> res(launches.size());
        check(garg_probe(s.ctx, nullptr, launches.data(), (uint32_t)launches.size(), &pu, s.tables,
                         GARG_INTEGRATOR_REFERENCE, res.data()),
              "probing");
        std::vector<double> drifts;
        int exhausted = 0;
        for (const auto& r : res) {
            drifts.push_back(r.drift);
            exhausted += r.outcome == GARG_RAY_EXHAUSTED;
        }
        std::sort(drifts.begin(), drifts.end());
        const double p99 = drifts[std::min((size_t)((drifts.size() - 1) * 0.99), drifts.size() - 1)];
        const std::vector<double> image = render(p.g);
        double difference = 0, total = 0;
        for (size_t i = 0; i < image.size(); ++i) {
            difference += std::fabs(spec[i] - image[i]);
            total += spec[i];
        }
        snprintf(buf, sizeof buf, "%-38s  %8.3f  %9.2e  %9.2e  %9d  %9.2e", p.name.c_str(), tm.mean, drifts.back(), p99,
                 exhausted, total > 0 ? difference / total : 0.0);
        line(buf);
    }
    line("splat phase breakdown");
    uint32_t frame = 0;
    Timing bin = time_pass(s.ctx, [&] { check(garg_splat_bin(s.volume, s.particles, nullptr, 0u, frame++), "splatting"); },
                           15);
    snprintf(buf, sizeof buf, "bin      %7.3f ms", bin.mean);
    line(buf);
    Timing resolve = time_pass(s.ctx, [&] { check(garg_splat_resolve(s.volume, nullptr, 1u), "resolving"); }, 15);
    snprintf(buf, sizeof buf, "resolve  %7.3f ms", resolve.mean);
    line(buf);
    line("launch shape sweep at the spec policy (batched march: threads per workgroup, resident waves per SIMD; "
         "one 8 x 8 tile per wave64)");
    garg_march_uniforms u = uniforms_for(policies[2].g, true);
    garg_march_request r{};
    r.uniforms = u;
    r.targets = t;
    r.tables = s.tables;
    r.volume = s.volume;
    for (uint32_t block : {64u, 128u, 256u})
        for (uint32_t waves : {4u, 5u, 8u}) {
            garg_batch_config bc;
            check(garg_batch_config_default(s.ctx, &bc), "reading the batch configuration");
            bc.block = block;
            bc.waves_per_simd = waves;
            Timing tm = time_pass(s.ctx, [&] { check(garg_march_batch(s.ctx, nullptr, &r, 1, &bc, nullptr), "marching"); },
                                  15);
            snprintf(buf, sizeof buf, "%3u threads, %u waves/SIMD  march %7.3f ms", block, waves, tm.mean);
            line(buf);
        }
    Timing plain = time_pass(
        s.ctx,
        [&] { check(garg_march_image(s.ctx, nullptr, &u, s.volume, s.tables, t, 0, 0, GARG_INTEGRATOR_REFERENCE), "marching"); },
        15);
    snprintf(buf, sizeof buf, "plain launch, 64 threads (8 x 8)  march %7.3f ms", plain.mean);
    line(buf);
    garg_targets_destroy(t);
}

// `bench --sweep-batch`: the batch autotuner on the headline frame (1920 x
// 1080, spin 0.9, 4096-step cap, observer r = 50, inclination 60 deg, 60 deg
// field of view) in the polar form: every block size x resident waves x
// compaction period x refill threshold (threshold 0, whole-tile claims, once
// per block and wave count). The fastest is stored for this device as the
// "headline" workload in the table that gargantua_amd.autotune.best() reads:
// <autotune_dir>/<arch>/headline.json = {"config": {...}, "ms": ..., "date": ...}.
static void sweep_batch(Scene& s) {
    OrbitCamera cam;
    cam.distance = 50.0f;
    cam.elevation = (float)(30.0 * M_PI / 180.0);
    cam.azimuth = 0.0f;
    cam.fov_y = (float)(60.0 * M_PI / 180.0);
    garg_march_uniforms u = make_uniforms(cam, 1920, 1080, Policy::still, true, 1u, true, 0.9, true);
    u.geodesic.escape_radius = 60.0f;
    garg_targets* t = nullptr;
    check(garg_targets_create(s.ctx, 1920, 1080, &t), "allocating targets");
    garg_march_request r{};
    r.uniforms = u;
    r.targets = t;
    r.volume = s.volume;
    garg_batch_config best{};
    double best_ms = 1e30;
    line("batch autotuner: 1920x1080, spin 0.9, 4096-step cap, polar form; block, waves/SIMD, period, threshold, ms");
    for (uint32_t block : {64u, 128u, 256u})
        for (uint32_t waves : {4u, 5u, 8u})
            for (uint32_t threshold : {0u, 16u, 32u, 48u, 63u})
                for (uint32_t period : {16u, 32u, 64u, 128u}) {
                    if (threshold == 0 && period != 16u) continue;  // whole tiles: the period plays no part
                    garg_batch_config c{block, waves, period, threshold, 2u};
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
    garg_device_info info;
    check(garg_device_info_get(s.ctx, &info), "reading the device");
    const std::string dir = autotune_dir() + "/" + info.arch;
    for (size_t k = 0; k <= dir.size(); ++k)
        if (k == dir.size() || (dir[k] == '/' && k > 0)) mkdir(dir.substr(0, k).c_str(), 0755);
    const std::string path = dir + "/headline.json";
    time_t now = time(nullptr);
    char date[16];
    strftime(date, sizeof date, "%Y-%m-%d", gmtime(&now));
    FILE* f = fopen(path.c_str(), "w");
    if (!f) operational("Could not write the autotuner's table " + path + ".");
    fprintf(f,
            "{\"config\": {\"block\": %u, \"compaction_period\": %u, \"kerr_schild\": %u, \"refill_threshold\": %u, "
            "\"waves_per_simd\": %u}, \"date\": \"%s\", \"ms\": %.4f}\n",
            best.block, best.compaction_period, best.kerr_schild, best.refill_threshold, best.waves_per_simd, date,
            best_ms);
    fclose(f);
    char buf[256];
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
    if (o.sweep_batch) {
        sweep_batch(s);
        s.close();
        exit(0);
    }
    if (o.sweep) {
        check(garg_splat_bin(s.volume, s.particles, nullptr, 0u, 0u), "splatting");
        check(garg_splat_resolve(s.volume, nullptr, 1u), "resolving");
        sweep(s, c);
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
    struct utsname uts;
    std::string os = uname(&uts) == 0 ? std::string(uts.sysname) + " " + uts.release : std::string("unknown");
    fprintf(f, "{\n  \"bake\": {\"bytes\": %llu, \"overflow_rays\": %u},\n  \"date\": \"%s\",\n  \"device\": \"%s %s\",\n",
            (unsigned long long)bake_bytes, overflow, stamp, info.name, info.arch);
    fprintf(f, "  \"frames\": {\"bake_ms\": %.4f, \"march_ms\": %.4f},\n  \"iterations\": %d,\n", bake_frame,
            march_frame, kIterations);
    // No upscaler: the presented image is the march's own resolution.
    fprintf(f, "  \"march\": [%d, %d],\n  \"os\": \"%s\",\n  \"output\": [%d, %d],\n  \"particles\": %d,\n  \"passes\": {\n",
            w, h, os.c_str(), w, h, c.particles);
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
