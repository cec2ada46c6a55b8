// `gargantua serve`: a continuous-batching render server (no reference
// counterpart; the nearest is the 3-in-flight interactive loop,
// App/Renderer.swift:43,119-215, and FrameStats' p50/p99, Stats.swift:48-68).
//
// Requests arrive as a Poisson stream. Each is a still or an orbit sequence
// with a resolution, spin, inclination, observer radius and step budget drawn
// from a seeded generator. A request is admitted, in arrival order, when its
// footprint (targets, plus bake slabs for static Schwarzschild stills) fits
// the HBM budget left beside the resident disk state and the batch has a
// slot. Every tick the disk advances one frame into a fresh volume version,
// and one persistent batched launch marches a bounded slice of rays from
// every admitted request, so requests that arrived while others were on the
// GPU join the next tick and share the device with them. A request pins the
// newest volume version at admission and renders all its frames against it,
// which makes its result independent of what it was batched with; --verify
// re-renders requests alone (march, or bake and walk) and compares the bits.
// With --max-wait, a request whose predicted wait (the admitted work left
// plus its own, in RK4 steps, over the measured rate) exceeds the bound is
// shed at admission instead of queueing behind more work than the GPU can
// finish in time.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gargantua_constants.h"
#include "host.hpp"

namespace host {

namespace {

using Clock = std::chrono::steady_clock;

struct Resolution {
    int w, h;
};
const Resolution kResolutions[] = {{256, 144}, {640, 360}, {960, 540}, {1920, 1080}, {3840, 2160}};

struct VolumeVersion {
    garg_volume* volume = nullptr;
    int pins = 0;
    uint64_t version = 0;
};

struct Request {
    int id = 0;
    double arrival = 0;  // seconds after start
    bool sequence = false;
    bool bake = false;
    int w = 0, h = 0, frames = 1;
    double spin = 0, inclination = 0, r_obs = 0;
    uint32_t budget = 256;
    // runtime
    double admitted = -1, done = -1;
    int frame = 0;
    int frame_slices = 0;  // batched launches the current frame's rays were split over
    uint32_t rays = 0, issued = 0, tick_issued = 0;
    int volume_slot = -1;
    uint64_t footprint = 0;
    garg_targets* targets = nullptr;
    garg_bake* baked = nullptr;
    bool baked_ready = false;
    bool shed = false;
    garg_march_uniforms uniforms{};
    uint64_t steps = 0, exhausted = 0, truncated = 0;
};

uint64_t fnv(const void* p, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
    const uint8_t* b = (const uint8_t*)p;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

uint64_t target_bytes(int w, int h) { return (uint64_t)w * h * 24u + 4096u; }

// Camera of frame f: an orbit sequence turns the azimuth once over its frames.
garg_march_uniforms frame_uniforms(const Request& r) {
    OrbitCamera cam;
    cam.azimuth = r.sequence ? (float)(2.0 * M_PI * r.frame / r.frames) : 0.6f;
    cam.elevation = (float)((90.0 - r.inclination) * M_PI / 180.0);
    cam.distance = (float)r.r_obs;
    cam.fov_y = (float)(60.0 * M_PI / 180.0);
    garg_march_uniforms u = make_uniforms(cam, r.w, r.h, r.budget == STEP_CAP_STILL ? Policy::still : Policy::interactive,
                                          true, 1u, true, r.spin, r.budget == STEP_CAP_STILL);
    u.geodesic.escape_radius = std::max(GARG_F(R_ESCAPE), (float)r.r_obs + 10.0f);
    return u;
}

// The walk takes the request's own uniforms with the bake spacing; the bake
// the still policy and no jitter, as the reference bakes
// (App/Renderer.swift:171-176), keeping the request's escape radius.
void bake_uniforms(const Request& r, garg_tables* t, int volume_size, garg_march_uniforms& walk,
                   garg_march_uniforms& bake) {
    walk = r.uniforms;
    check(garg_tables_apply(t, &walk), "applying tables");
    walk.bake_spacing = (float)(2.0 * VOLUME_HALF_EXTENT / volume_size * BAKE_SPACING_VOXELS);
    bake = walk;
    bake.geodesic = policy_settings(Policy::still);
    bake.geodesic.escape_radius = walk.geodesic.escape_radius;
    bake.jitter[0] = bake.jitter[1] = 0.0f;
}

}  // namespace

void run_serve(const Options& o) {
    Configuration c = Configuration::make(o);
    // The disk: the preset's particles and volume, advanced every tick.
    Scene s = make_scene(c, (int)STILL_SETTLE_STEPS, true);
    garg_device_info info;
    check(garg_device_info_get(s.ctx, &info), "reading the device");
    const int kVersions = 3;
    std::vector<VolumeVersion> versions(kVersions);
    versions[0].volume = s.volume;
    for (int i = 1; i < kVersions; ++i)
        check(garg_volume_create(s.ctx, (uint32_t)c.volume, GARG_F(VOLUME_HALF_EXTENT), 0.0f, &versions[i].volume),
              "allocating a volume version");
    int newest = 0;
    uint64_t version_counter = 0;
    // One captured graph per version: a sim dispatch, the stochastic splat
    // (its frame is the dispatch count, so every replay draws a new one) and
    // the resolve, replayed each tick instead of re-enqueueing three launches.
    std::vector<garg_graph*> disk_frame(kVersions, nullptr);
    for (int i = 0; i < kVersions; ++i) {
        check(garg_graph_begin(s.ctx, nullptr), "capturing the disk frame");
        check(garg_particles_step(s.particles, nullptr, 1), "capturing the disk frame");
        check(garg_splat_bin(versions[i].volume, s.particles, nullptr, 0u, GARG_FRAME_AUTO), "capturing the disk frame");
        check(garg_splat_resolve(versions[i].volume, nullptr, 1u), "capturing the disk frame");
        check(garg_graph_end(s.ctx, nullptr, &disk_frame[i]), "capturing the disk frame");
    }
    // Sweep tables per observer radius for Schwarzschild requests.
    struct TableSet {
        float radius;
        garg_tables* t;
    };
    std::vector<TableSet> tables;
    auto tables_for = [&](float radius) -> garg_tables* {
        for (auto& t : tables)
            if (t.radius == radius) return t.t;
        garg_tables* t = nullptr;
        check(garg_tables_create(s.ctx, nullptr, GARG_F(VOLUME_HALF_EXTENT), &t), "building tables");
        check(garg_tables_update(t, nullptr, radius), "building tables");
        tables.push_back({radius, t});
        return t;
    };

    // Request stream: Poisson arrivals at `rate` for `seconds`.
    std::mt19937_64 rng(o.serve_seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<Request> all;
    for (double t = 0;;) {
        t += -std::log(1.0 - U(rng)) / o.serve_rate;
        if (t > o.serve_seconds) break;
        Request r;
        r.id = (int)all.size();
        r.arrival = t;
        r.sequence = U(rng) < 0.5;
        Resolution res = kResolutions[std::min(4, (int)(U(rng) * 5))];
        r.w = res.w;
        r.h = res.h;
        r.spin = 0.99 * U(rng);
        r.inclination = 10.0 + 79.0 * U(rng);
        r.r_obs = 30.0 + 70.0 * U(rng);
        r.budget = U(rng) < 0.5 ? STEP_CAP_INTERACTIVE : STEP_CAP_STILL;
        r.frames = r.sequence ? o.serve_sequence_frames : 1;
        // With the bake strategy a stated share of the stills are static
        // Schwarzschild stills, and these bake: the bake is Schwarzschild only
        // (the reference refuses spin with it, App/Renderer.swift:62-64).
        if (!r.sequence && c.strategy == Strategy::bake && U(rng) < o.serve_bake_share) r.spin = 0.0;
        r.bake = !r.sequence && c.strategy == Strategy::bake && r.spin == 0.0;
        r.rays = garg_request_rays((uint32_t)r.w, (uint32_t)r.h);
        r.footprint = target_bytes(r.w, r.h) + (r.bake ? garg_bake_bytes(r.w, r.h, BAKE_SAMPLES) : 0);
        all.push_back(r);
    }
    uint64_t resident = 0;
    check(garg_ctx_allocated(s.ctx, &resident), "reading allocations");
    const uint64_t serving_budget = info.budget_bytes > resident ? info.budget_bytes - resident : 0;
    line("serving " + std::to_string(all.size()) + " requests over " + std::to_string((int)o.serve_seconds) +
         " s at " + std::to_string(o.serve_rate) + " req/s; budget for requests " + format_bytes(serving_budget));

    garg_batch_config cfg;
    check(garg_batch_config_default(s.ctx, &cfg), "reading the batch configuration");
    cfg.kerr_schild = o.form;
    const uint64_t kTickRays = (uint64_t)o.serve_tick_rays;  // rays marched per tick across all requests
    size_t next_arrival = 0;
    std::deque<int> pending;
    std::vector<int> active;
    uint64_t in_use = 0, total_steps = 0, ticks = 0, refused = 0, shed = 0, verified = 0, mismatched = 0;
    uint64_t verified_bakes = 0;
    // Expected RK4 steps per ray (interactive, still budget), refined from
    // the frames that finish, and the marching rate, for --max-wait.
    double steps_per_ray[2] = {100.0, 200.0};
    uint64_t seen_rays[2] = {0, 0}, seen_steps[2] = {0, 0};
    double march_seconds = 0;
    auto kind = [](const Request& r) { return r.budget > STEP_CAP_INTERACTIVE ? 1 : 0; };
    auto work_left = [&](const Request& r) {
        const uint64_t rays = r.bake ? (r.baked_ready ? 0 : r.rays)
                                     : (uint64_t)(r.frames - r.frame - 1) * r.rays + (r.rays - r.issued);
        return (double)rays * steps_per_ray[r.bake ? 1 : kind(r)];
    };
    double gpu_busy = 0;
    const auto start = Clock::now();
    auto now_s = [&] { return std::chrono::duration<double>(Clock::now() - start).count(); };
    size_t completed = 0;

    double next_progress = 20.0;
    while (completed + refused + shed < all.size()) {
        double now = now_s();
        if (o.serve_deadline > 0 && now > o.serve_deadline) break;  // report what finished
        if (now >= next_progress) {
            line("t=" + std::to_string((int)now) + " s: " + std::to_string(completed) + " done, " +
                 std::to_string(active.size()) + " on the GPU, " + std::to_string(pending.size()) + " waiting, " +
                 std::to_string(all.size() - next_arrival) + " yet to arrive");
            next_progress += 20.0;
        }
        while (next_arrival < all.size() && all[next_arrival].arrival <= now) pending.push_back((int)next_arrival++);
        if (pending.empty() && active.empty()) {
            // Idle until the next arrival.
            double wait = all[next_arrival].arrival - now;
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            continue;
        }
        // Admission in arrival order by HBM footprint and batch slots.
        while (!pending.empty()) {
            Request& r = all[pending.front()];
            if (r.footprint > serving_budget) {  // can never fit: refuse
                r.done = now;
                ++refused;
                pending.pop_front();
                continue;
            }
            if (o.serve_max_wait > 0) {
                // Held back by the HBM budget or the batch's slots for longer
                // than the wait: shed, so the queue stays bounded too.
                if (now - r.arrival > o.serve_max_wait) {
                    r.done = now;
                    r.shed = true;
                    ++shed;
                    pending.pop_front();
                    continue;
                }
                double ahead = work_left(r);
                for (int id : active) ahead += work_left(all[id]);
                const double rate = march_seconds > 0.5 ? total_steps / march_seconds : 5e10;
                if (ahead / rate > o.serve_max_wait) {
                    r.done = now;
                    r.shed = true;
                    ++shed;
                    pending.pop_front();
                    continue;
                }
            }
            if (in_use + r.footprint > serving_budget || active.size() >= 200) break;
            check(garg_targets_create(s.ctx, (uint32_t)r.w, (uint32_t)r.h, &r.targets), "allocating targets");
            if (r.bake) check(garg_bake_create(s.ctx, (uint32_t)r.w, (uint32_t)r.h, BAKE_SAMPLES, &r.baked), "baking");
            r.admitted = now;
            r.volume_slot = newest;
            versions[newest].pins += 1;
            in_use += r.footprint;
            r.uniforms = frame_uniforms(r);
            active.push_back(r.id);
            pending.pop_front();
        }
        // Advance the disk one frame into a free volume version.
        int slot = -1;
        for (int k = 1; k <= kVersions; ++k) {
            int cand = (newest + k) % kVersions;
            if (cand != newest && versions[cand].pins == 0) {
                slot = cand;
                break;
            }
        }
        if (slot >= 0) {
            check(garg_graph_launch(disk_frame[slot], nullptr), "advancing the disk");
            versions[slot].version = ++version_counter;
            newest = slot;
        }
        // One batched launch over a slice of every active request's current frame.
        std::vector<garg_march_request> batch;
        uint64_t budget = kTickRays;
        // Every active request starts the tick with nothing issued, also those
        // the tick's ray budget does not reach.
        for (int id : active) all[id].tick_issued = 0;
        // The order the budget is handed out in: shortest remaining work
        // first (rays left over all frames, a still-budget ray counted as two
        // since it takes about twice the steps; ties in arrival order), or
        // arrival order. Ray results do not depend on it. Shortest first
        // gives the lowest median and p99 on the serving stream; its cost is
        // the tail: a long sequence gets what newer, shorter arrivals leave.
        // (Serving requests that have waited longest first, or giving the
        // oldest an eighth of each tick, or ageing the order, bounded that
        // tail less well than it cost the p99: the admission estimate of a
        // request's own work can be several times short for long sequences
        // full of exhausted rays.)
        std::vector<int> order(active);
        if (o.serve_policy == "srpt") {
            auto work = [&](int id) {
                const Request& r = all[id];
                const uint64_t left = (uint64_t)(r.frames - r.frame - 1) * r.rays + (r.rays - r.issued);
                return r.budget > STEP_CAP_INTERACTIVE ? 2 * left : left;
            };
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) < work(y); });
        }
        for (int id : order) {
            Request& r = all[id];
            if (r.bake) continue;
            if (budget == 0) break;
            uint32_t take = (uint32_t)std::min<uint64_t>(budget, r.rays - r.issued);
            if (take == 0) continue;
            garg_march_request q{};
            q.uniforms = r.uniforms;
            q.targets = r.targets;
            q.volume = versions[r.volume_slot].volume;
            if (r.spin == 0.0) {
                q.tables = tables_for((float)r.r_obs);
                check(garg_tables_apply(q.tables, &q.uniforms), "applying tables");
            }
            q.ray_begin = r.issued;
            q.ray_count = take;
            batch.push_back(q);
            r.frame_slices += 1;
            r.tick_issued = take;
            budget -= take;
        }
        auto t0 = Clock::now();
        if (!batch.empty()) {
            garg_march_counters total;
            check(garg_march_batch(s.ctx, nullptr, batch.data(), (uint32_t)batch.size(), &cfg, &total), "marching");
            total_steps += total.steps;
            march_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
        }
        // Bake requests: bake once, then walk against their pinned volume.
        for (int id : active) {
            Request& r = all[id];
            if (!r.bake) continue;
            garg_tables* t = tables_for((float)r.r_obs);
            garg_march_uniforms walk, bake;
            bake_uniforms(r, t, c.volume, walk, bake);
            if (!r.baked_ready) {
                check(garg_bake_rows(r.baked, nullptr, &bake, t, 0, (uint32_t)r.h), "baking");
                r.baked_ready = true;
            }
            check(garg_bake_walk(r.baked, nullptr, &walk, versions[r.volume_slot].volume, r.targets), "walking");
            r.tick_issued = r.rays - r.issued;
        }
        check(garg_stream_sync(s.ctx, nullptr), "waiting for the tick");
        gpu_busy += std::chrono::duration<double>(Clock::now() - t0).count();
        ++ticks;
        // Retire frames and requests.
        now = now_s();
        std::vector<int> still_active;
        for (int id : active) {
            Request& r = all[id];
            r.issued += r.tick_issued;
            if (r.issued < r.rays) {
                still_active.push_back(id);
                continue;
            }
            garg_march_counters cnt;
            check(garg_targets_read_counters(r.targets, &cnt), "reading counters");
            r.exhausted = cnt.exhausted;
            if (!r.bake && cnt.rays > 0) {
                // The counters accumulate over the request's frames.
                const int k = kind(r);
                seen_rays[k] += r.rays;
                seen_steps[k] += cnt.steps - r.steps;
                r.steps = cnt.steps;
                steps_per_ray[k] = (double)seen_steps[k] / (double)seen_rays[k];
            }
            if (r.baked) {
                uint32_t over = 0;
                check(garg_bake_overflow(r.baked, &over), "reading the bake overflow");
                r.truncated = over;
            }
            if (++r.frame < r.frames) {
                r.issued = 0;
                r.frame_slices = 0;
                r.uniforms = frame_uniforms(r);
                still_active.push_back(id);
                continue;
            }
            // Finished: verify alone (every Nth request), release.
            if (o.serve_verify > 0 && r.id % o.serve_verify == 0) {
                const size_t n = (size_t)r.w * r.h;
                std::vector<uint16_t> mixed(n * 8), alone(n * 8);
                std::vector<float> mixed_depth(n), alone_depth(n);
                check(garg_targets_read(r.targets, mixed.data(), mixed.data() + n * 4, mixed_depth.data()), "reading");
                garg_targets* t2 = nullptr;
                check(garg_targets_create(s.ctx, (uint32_t)r.w, (uint32_t)r.h, &t2), "allocating");
                if (r.bake) {
                    garg_tables* t = tables_for((float)r.r_obs);
                    garg_march_uniforms walk, bake;
                    bake_uniforms(r, t, c.volume, walk, bake);
                    garg_bake* b2 = nullptr;
                    check(garg_bake_create(s.ctx, (uint32_t)r.w, (uint32_t)r.h, BAKE_SAMPLES, &b2), "baking alone");
                    check(garg_bake_rows(b2, nullptr, &bake, t, 0, (uint32_t)r.h), "baking alone");
                    check(garg_bake_walk(b2, nullptr, &walk, versions[r.volume_slot].volume, t2), "walking alone");
                    check(garg_stream_sync(s.ctx, nullptr), "waiting for the bake");
                    garg_bake_destroy(b2);
                } else {
                    garg_march_request q{};
                    q.uniforms = r.uniforms;
                    q.targets = t2;
                    q.volume = versions[r.volume_slot].volume;
                    if (r.spin == 0.0) {
                        q.tables = tables_for((float)r.r_obs);
                        check(garg_tables_apply(q.tables, &q.uniforms), "applying tables");
                    }
                    check(garg_march_batch(s.ctx, nullptr, &q, 1, &cfg, nullptr), "marching alone");
                }
                check(garg_targets_read(t2, alone.data(), alone.data() + n * 4, alone_depth.data()), "reading");
                garg_targets_destroy(t2);
                ++verified;
                if (r.bake) ++verified_bakes;
                if (fnv(mixed.data(), mixed.size() * 2, fnv(mixed_depth.data(), n * 4)) !=
                    fnv(alone.data(), alone.size() * 2, fnv(alone_depth.data(), n * 4))) {
                    ++mismatched;
                    size_t differing = 0, first = 0;
                    for (size_t k = 0; k < n; ++k)
                        if (memcmp(&mixed[4 * k], &alone[4 * k], 8) != 0 ||
                            memcmp(&mixed[n * 4 + 4 * k], &alone[n * 4 + 4 * k], 8) != 0 ||
                            memcmp(&mixed_depth[k], &alone_depth[k], 4) != 0)
                            if (differing++ == 0) first = k;
                    char msg[512];
                    snprintf(msg, sizeof msg,
                             "mismatch: request %d %dx%d spin %.3f budget %u frames %d slices %d slot %d%s: %zu pixels "
                             "differ, first (%zu, %zu)",
                             r.id, r.w, r.h, r.spin, r.budget, r.frames, r.frame_slices, r.volume_slot,
                             r.bake ? " (baked)" : "", differing, first % (size_t)r.w, first / (size_t)r.w);
                    line(msg);
                }
            }
            r.done = now_s();
            versions[r.volume_slot].pins -= 1;
            in_use -= r.footprint;
            garg_targets_destroy(r.targets);
            garg_bake_destroy(r.baked);
            r.targets = nullptr;
            r.baked = nullptr;
            ++completed;
        }
        active.swap(still_active);
    }
    double wall = now_s();
    std::vector<double> latency;
    uint64_t exhausted = 0, truncated = 0, baked = 0;
    size_t unfinished = 0;
    for (auto& r : all) {
        if (r.done < 0) ++unfinished;  // past the deadline: still queued or on the GPU
        if (r.bake && r.done >= 0 && !r.shed && r.admitted >= 0) ++baked;
        if (r.admitted >= 0 && r.done >= 0) {
            latency.push_back(r.done - r.arrival);
            exhausted += r.exhausted;
            truncated += r.truncated;
        }
    }
    std::sort(latency.begin(), latency.end());
    auto pct = [&](double p) {
        return latency.empty() ? 0.0 : latency[std::min(latency.size() - 1, (size_t)((latency.size() - 1) * p))];
    };
    char buf[1024];
    snprintf(buf, sizeof buf,
             "{\"requests\": %zu, \"completed\": %zu, \"unfinished\": %zu, \"refused\": %llu, \"shed\": %llu, "
             "\"max_wait_s\": %.3f, \"baked\": %llu, \"wall_s\": %.3f, "
             "\"ticks\": %llu, "
             "\"latency_p50_s\": %.4f, \"latency_p99_s\": %.4f, \"latency_max_s\": %.4f, \"steps\": %llu, "
             "\"steps_per_s\": %.4e, \"gpu_busy_fraction\": %.3f, \"exhausted_rays\": %llu, \"truncated_rays\": %llu, "
             "\"verified\": %llu, \"verified_baked\": %llu, \"mismatched\": %llu, \"rate\": %.3f, "
             "\"arrival_seconds\": %.3f, \"frames\": %d, \"policy\": \"%s\", \"form\": \"%s\"}",
             all.size(), completed, unfinished, (unsigned long long)refused, (unsigned long long)shed, o.serve_max_wait,
             (unsigned long long)baked, wall, (unsigned long long)ticks, pct(0.5),
             pct(0.99),
             latency.empty() ? 0.0 : latency.back(), (unsigned long long)total_steps, total_steps / wall,
             gpu_busy / wall, (unsigned long long)exhausted, (unsigned long long)truncated,
             (unsigned long long)verified, (unsigned long long)verified_bakes, (unsigned long long)mismatched,
             o.serve_rate, o.serve_seconds,
             o.serve_sequence_frames, o.serve_policy.c_str(), form_name(o.form));
    line(buf);
    if (!o.serve_report.empty()) {
        FILE* f = fopen(o.serve_report.c_str(), "w");
        if (f) {
            fprintf(f, "%s\n", buf);
            fclose(f);
        }
    }
    for (auto& t : tables) garg_tables_destroy(t.t);
    for (garg_graph* g : disk_frame) garg_graph_destroy(g);
    for (int i = 1; i < kVersions; ++i) garg_volume_destroy(versions[i].volume);
    s.close();
    exit(mismatched == 0 ? 0 : 1);
}

}  // namespace host
