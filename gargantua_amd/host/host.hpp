// The C++ host of the gargantua CLI, above the C ABI of libgargantua_hip.so:
// the option grammar and exit codes (Options.swift, Exit.swift), presets and
// the memory budget (Presets.swift, Budget.swift), the orbit camera and the
// march uniforms (Camera.swift, MarchPass.swift:61-98), and the commands.
#pragma once

#include <cstdint>
#include <optional>
#include <string>
#include <vector>

#include "../../include/gargantua.h"

namespace host {

// ---- exits (Exit.swift:5-17) ----------------------------------------------
[[noreturn]] void operational(const std::string& sentence);  // exit 2
[[noreturn]] void usage_error(const std::string& sentence);   // exit 64
void line(const std::string& text);                           // stdout, flushed
// Exit 2 with the library's sentence when a call fails.
void check(int status, const char* what);

// ---- options (Options.swift:3-147) -----------------------------------------
enum class Strategy { march, bake };
enum class Preset { small, standard, max };
enum class Command { interactive, still, bench, validate, version, constants, serve };

struct Options {
    Command command = Command::interactive;
    std::string still_path;
    Preset preset = Preset::standard;
    std::optional<int> particles;
    std::optional<int> volume;
    uint64_t seed = 1;
    Strategy strategy = Strategy::march;
    int fps_cap = 60;
    bool upscale = true;
    int width = 3840;
    int height = 2160;
    std::optional<double> soak_seconds;
    bool sweep = false;        // bench: the reference's step-policy, splat and launch-shape sweep
    bool sweep_batch = false;  // bench: the batch autotuner over the persistent march's launch
    std::optional<int> sequence;
    std::optional<double> cam_azimuth, cam_elevation, cam_distance;
    double spin = 0.0;
    // How spinning rays step (garg_batch_config.kerr_schild): 0 Boyer-Lindquist
    // fast steps throughout (the reference's coordinates; its E/L drift blows up
    // near the spin axis), 1 Kerr-Schild coordinates throughout (regular on the
    // axis), 2 the polar form (Boyer-Lindquist until a ray's sin(theta) drops
    // below 0.25, Kerr-Schild from there), which holds E/L drift under 1e-5 on
    // escaping rays at the cost of the fast form plus a few percent.
    unsigned form = 2;
    // serve: arrival rate, duration, scale-down of the workload for a bounded run
    double serve_rate = 20.0;
    double serve_seconds = 300.0;
    double serve_deadline = 0.0;  // stop the loop after this wall time (0 = run until every request is done)
    int serve_tick_rays = 4 << 20;  // rays marched per tick across all requests
    std::string serve_report;
    uint64_t serve_seed = 1;
    int serve_sequence_frames = 240;
    int serve_verify = 0;  // re-render every Nth request alone and compare bits
    // How a tick's ray budget is shared: "srpt" gives it to the active
    // requests with the least work left first, "fifo" in arrival order.
    std::string serve_policy = "srpt";
    // With --strategy bake: the share of still requests that are static
    // Schwarzschild stills (spin 0), which bake once and walk.
    double serve_bake_share = 0.25;
    // Shed a request at admission when the work already admitted plus its
    // own, at the measured rate, would take longer than this (0 = never).
    double serve_max_wait = 0.0;
};

extern const char* kUsage;
// Pure: no device access, so it can be exercised without a GPU.
// Returns an empty string on success, else the usage sentence.
std::string parse_options(const std::vector<std::string>& args, Options& out);

// ---- presets and budget (Presets.swift, Budget.swift) ----------------------
struct Configuration {
    Preset preset;
    int particles;
    int volume;
    int march_width;
    int march_height;
    uint64_t seed;
    double spin;
    Strategy strategy;
    static Configuration make(const Options& o);
    uint64_t bytes() const;
};
const char* preset_name(Preset p);
const char* form_name(unsigned form);  // the --form spelling of Options::form
std::string format_bytes(uint64_t bytes);
// One sentence explaining a refusal, or empty when `bytes` fit the budget.
std::string budget_refusal(uint64_t bytes, uint64_t budget, const std::string& what);

// ---- camera and uniforms (Camera.swift:5-33, MarchPass.swift:61-98) ------
struct OrbitCamera {
    float azimuth = 0.6f;
    float elevation = 0.32f;
    float distance = 52.0f;
    float fov_y = 50.0f * 3.14159265358979f / 180.0f;
    static constexpr float kMinimumDistance = 6.0f;
    static constexpr float kMaximumDistance = 56.0f;
    void basis(float position[3], float right[3], float up[3], float forward[3]) const;
};
OrbitCamera camera_from(const Options& o);

// Probes.cameraGrid (Validation/Probes.swift:52-66).
std::vector<garg_ray_probe> camera_grid(const OrbitCamera& cam, int columns, int rows, double aspect);
// The batch autotuner's table directory (GARG_AUTOTUNE_DIR, else beside the library).
std::string autotune_dir();

enum class Policy { interactive, still };
garg_geodesic_settings policy_settings(Policy p);
garg_march_uniforms make_uniforms(const OrbitCamera& camera, int width, int height, Policy policy, bool redshift,
                                  uint32_t star_seed, bool stars, double spin, bool compensated);

// ---- PNG (App/Capture.swift:72-94 writes through CoreGraphics) --------------
// 8-bit RGBA, stored (uncompressed) deflate blocks, CRC-32 and Adler-32.
bool write_png(const std::string& path, const uint8_t* rgba, int width, int height);
// IEEE binary16 to float.
float half_to_float(uint16_t h);

// ---- the scene every command starts from ------------------------------------
struct Scene {
    garg_ctx* ctx = nullptr;
    garg_particles* particles = nullptr;
    garg_volume* volume = nullptr;
    garg_tables* tables = nullptr;
    void close();
};
// Context, particles seeded and settled `settle_substeps`, volume resolved.
Scene make_scene(const Configuration& c, int settle_substeps, bool exact_splat);

// ---- commands -----------------------------------------------------------------
[[noreturn]] void run_still(const Options& o);
[[noreturn]] void run_bench(const Options& o);
[[noreturn]] void run_validate(const Options& o);
[[noreturn]] void run_serve(const Options& o);
void print_constants();

}  // namespace host
