// Host pieces shared by the commands: exits, the option grammar, presets and
// budget, the orbit camera and uniforms, the PNG writer and the scene setup.
#include <cmath>
#include <dlfcn.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gargantua_constants.h"
#include "host.hpp"

namespace host {

void operational(const std::string& sentence) {
    fprintf(stderr, "%s\n", sentence.c_str());
    exit(2);
}

void usage_error(const std::string& sentence) {
    fprintf(stderr, "%s\n", sentence.c_str());
    exit(64);
}

void line(const std::string& text) {
    fprintf(stdout, "%s\n", text.c_str());
    fflush(stdout);
}

void check(int status, const char* what) {
    if (status != GARG_OK) operational(std::string(what) + ": " + garg_last_error());
}

const char* kUsage =
    "usage: gargantua [--preset small|default|max] [--particles N] [--volume 96|128|192|256] [--seed S] "
    "[--strategy march|bake] [--fps-cap 30|60] [--no-upscale] [--spin A] [--form fast|kerr-schild|polar] [--camera az,el,dist] "
    "[--still out.png --width W --height H [--sequence N]] | bench [--sweep | --sweep-batch] | validate | serve [--rate R] "
    "[--seconds T] [--deadline D] [--frames F] [--tick-rays N] [--policy srpt|fifo] [--bake-share F] [--max-wait S] "
    "[--verify N] [--report PATH] | version | constants";

namespace {

bool parse_positive(const std::string& text, int& out) {
    char* end = nullptr;
    long v = strtol(text.c_str(), &end, 10);
    if (end == text.c_str() || *end != '\0' || v <= 0 || v > 1 << 30) return false;
    out = (int)v;
    return true;
}

bool parse_double(const std::string& text, double& out) {
    char* end = nullptr;
    out = strtod(text.c_str(), &end);
    return end != text.c_str() && *end == '\0' && std::isfinite(out);
}

}  // namespace

// Options.swift:64-147.
std::string parse_options(const std::vector<std::string>& args, Options& o) {
    size_t i = 0;
    auto value = [&](const std::string& flag, std::string& out) -> std::string {
        if (++i >= args.size()) return flag + " needs a value. " + kUsage;
        out = args[i];
        return "";
    };
    auto integer = [&](const std::string& flag, int& out) -> std::string {
        std::string text, err = value(flag, text);
        if (!err.empty()) return err;
        if (!parse_positive(text, out)) return flag + " needs a positive integer, not " + text + ".";
        return "";
    };
    for (; i < args.size(); ++i) {
        const std::string& a = args[i];
        std::string text, err;
        if (a == "bench") o.command = Command::bench;
        else if (a == "validate") o.command = Command::validate;
        else if (a == "version") o.command = Command::version;
        else if (a == "constants") o.command = Command::constants;
        else if (a == "serve") o.command = Command::serve;
        else if (a == "--preset") {
            if (!(err = value(a, text)).empty()) return err;
            if (text == "small") o.preset = Preset::small;
            else if (text == "default") o.preset = Preset::standard;
            else if (text == "max") o.preset = Preset::max;
            else return "unknown preset " + text + "; choose small, default or max.";
        } else if (a == "--particles") {
            int n;
            if (!(err = integer(a, n)).empty()) return err;
            o.particles = n;
        } else if (a == "--volume") {
            int n;
            if (!(err = integer(a, n)).empty()) return err;
            if (n != 96 && n != 128 && n != 192 && n != 256) return "volume must be 96, 128, 192 or 256.";
            o.volume = n;
        } else if (a == "--seed") {
            if (!(err = value(a, text)).empty()) return err;
            char* end = nullptr;
            unsigned long long s = strtoull(text.c_str(), &end, 10);
            if (end == text.c_str() || *end != '\0' || text[0] == '-')
                return "seed must be an unsigned 64 bit integer, not " + text + ".";
            o.seed = s;
        } else if (a == "--strategy") {
            if (!(err = value(a, text)).empty()) return err;
            if (text == "march") o.strategy = Strategy::march;
            else if (text == "bake") o.strategy = Strategy::bake;
            else return "unknown strategy " + text + "; choose march or bake.";
        } else if (a == "--fps-cap") {
            int n;
            if (!(err = integer(a, n)).empty()) return err;
            if (n != 30 && n != 60) return "fps cap must be 30 or 60.";
            o.fps_cap = n;
        } else if (a == "--no-upscale") {
            o.upscale = false;
        } else if (a == "--still") {
            if (!(err = value(a, text)).empty()) return err;
            o.command = Command::still;
            o.still_path = text;
        } else if (a == "--width") {
            if (!(err = integer(a, o.width)).empty()) return err;
        } else if (a == "--height") {
            if (!(err = integer(a, o.height)).empty()) return err;
        } else if (a == "--soak") {
            double s;
            if (!(err = value(a, text)).empty()) return err;
            if (!parse_double(text, s) || s <= 0) return "soak needs a positive number of seconds.";
            o.soak_seconds = s;
        } else if (a == "--sweep") {
            o.sweep = true;
        } else if (a == "--sweep-batch") {
            o.sweep_batch = true;
        } else if (a == "--sequence") {
            int n;
            if (!(err = integer(a, n)).empty()) return err;
            o.sequence = n;
        } else if (a == "--spin") {
            double s;
            if (!(err = value(a, text)).empty()) return err;
            if (!parse_double(text, s) || s < 0 || s > 0.998) return "spin must be between 0 and 0.998.";
            o.spin = s;
        } else if (a == "--kerr-schild") {
            o.form = 1;
        } else if (a == "--form") {
            if (!(err = value(a, text)).empty()) return err;
            if (text == "fast") o.form = 0;
            else if (text == "kerr-schild") o.form = 1;
            else if (text == "polar") o.form = 2;
            else return "form must be fast, kerr-schild or polar.";
        } else if (a == "--camera") {
            if (!(err = value(a, text)).empty()) return err;
            double p[3];
            int n = 0;
            size_t start = 0;
            bool ok = true;
            while (n < 4 && start <= text.size()) {
                size_t comma = text.find(',', start);
                std::string part = text.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
                if (n < 3 && !parse_double(part, p[n])) ok = false;
                ++n;
                if (comma == std::string::npos) break;
                start = comma + 1;
            }
            if (!ok || n != 3 || !(p[2] > 2.5 && p[2] < 60))
                return "camera needs azimuth,elevation,distance with distance between 2.5 and 60.";
            o.cam_azimuth = p[0];
            o.cam_elevation = p[1];
            o.cam_distance = p[2];
        } else if (a == "--quit") {
            if (!(err = value(a, text)).empty()) return err;
            if (text != "direct" && text != "esc" && text != "cmd-q")
                return "unknown quit path " + text + "; choose direct, esc or cmd-q.";
        } else if (a == "--rate") {
            if (!(err = value(a, text)).empty()) return err;
            if (!parse_double(text, o.serve_rate) || o.serve_rate <= 0) return "rate needs a positive number.";
        } else if (a == "--seconds") {
            if (!(err = value(a, text)).empty()) return err;
            if (!parse_double(text, o.serve_seconds) || o.serve_seconds <= 0) return "seconds needs a positive number.";
        } else if (a == "--deadline") {
            if (!(err = value(a, text)).empty()) return err;
            if (!parse_double(text, o.serve_deadline) || o.serve_deadline <= 0) return "deadline needs a positive number.";
        } else if (a == "--tick-rays") {
            if (!(err = integer(a, o.serve_tick_rays)).empty()) return err;
            if (o.serve_tick_rays < 64) return "tick-rays needs at least 64 rays.";
        } else if (a == "--frames") {
            if (!(err = integer(a, o.serve_sequence_frames)).empty()) return err;
        } else if (a == "--verify") {
            if (!(err = integer(a, o.serve_verify)).empty()) return err;
        } else if (a == "--policy") {
            if (!(err = value(a, o.serve_policy)).empty()) return err;
            if (o.serve_policy != "srpt" && o.serve_policy != "fifo")
                return "unknown policy " + o.serve_policy + "; choose srpt or fifo.";
        } else if (a == "--bake-share") {
            if (!(err = value(a, text)).empty()) return err;
            if (!parse_double(text, o.serve_bake_share) || o.serve_bake_share < 0 || o.serve_bake_share > 1)
                return "bake-share needs a number between 0 and 1.";
        } else if (a == "--max-wait") {
            if (!(err = value(a, text)).empty()) return err;
            if (!parse_double(text, o.serve_max_wait) || o.serve_max_wait <= 0) return "max-wait needs a positive number.";
        } else if (a == "--report") {
            if (!(err = value(a, o.serve_report)).empty()) return err;
        } else {
            return "unknown argument " + a + ". " + kUsage;
        }
    }
    return "";
}

const char* preset_name(Preset p) {
    switch (p) {
        case Preset::small: return "small";
        case Preset::standard: return "default";
        case Preset::max: return "max";
    }
    return "default";
}

const char* form_name(unsigned form) {
    return form == 0 ? "fast" : form == 1 ? "kerr-schild" : "polar";
}

// Presets.swift:18-31, 72-93.
Configuration Configuration::make(const Options& o) {
    Configuration c;
    c.preset = o.preset;
    int particles = o.preset == Preset::small ? 250000 : o.preset == Preset::standard ? 1000000 : 4000000;
    int volume = o.preset == Preset::small ? 96 : o.preset == Preset::standard ? 128 : 256;
    c.particles = o.particles ? *o.particles : particles;
    c.volume = o.volume ? *o.volume : volume;
    c.march_width = o.preset == Preset::small ? 640 : 960;
    c.march_height = o.preset == Preset::small ? 360 : 540;
    c.seed = o.seed;
    c.spin = o.spin;
    c.strategy = o.strategy;
    return c;
}

// Presets.swift:48-68 with this build's layouts: a 32-byte particle, a
// 24-byte sum and a 16-byte resolved voxel, 8 + 8 + 4 bytes per target
// pixel, and the bake slabs.
uint64_t Configuration::bytes() const {
    uint64_t voxels = (uint64_t)volume * volume * volume;
    uint64_t pixels = (uint64_t)march_width * march_height;
    uint64_t total = (uint64_t)particles * 32u + INVERSE_CDF_COUNT * 4u + BLACKBODY_TABLE_SIZE * 16u;
    total += voxels * 24u + voxels * 16u + pixels * 24u;
    if (strategy == Strategy::bake) total += garg_bake_bytes(march_width, march_height, BAKE_SAMPLES);
    return total;
}

std::string format_bytes(uint64_t bytes) {
    char buf[64];
    double v = (double)bytes;
    if (v >= 1e9) snprintf(buf, sizeof buf, "%.1f GB", v / 1e9);
    else if (v >= 1e6) snprintf(buf, sizeof buf, "%.0f MB", v / 1e6);
    else if (v >= 1e3) snprintf(buf, sizeof buf, "%.0f kB", v / 1e3);
    else snprintf(buf, sizeof buf, "%llu bytes", (unsigned long long)bytes);
    return buf;
}

// Budget.swift:16-20.
std::string budget_refusal(uint64_t bytes, uint64_t budget, const std::string& what) {
    if (bytes <= budget) return "";
    return what + " needs " + format_bytes(bytes) + " but this device's budget is " + format_bytes(budget) + ".";
}

void OrbitCamera::basis(float p[3], float r[3], float u[3], float f[3]) const {
    float c = std::cos(elevation);
    p[0] = distance * (c * std::cos(azimuth));
    p[1] = distance * (c * std::sin(azimuth));
    p[2] = distance * std::sin(elevation);
    float len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    for (int i = 0; i < 3; ++i) f[i] = -p[i] / len;
    // right = normalize(forward x z), up = right x forward (Camera.swift:31-33).
    float rx = f[1], ry = -f[0];
    float rl = std::sqrt(rx * rx + ry * ry);
    r[0] = rx / rl;
    r[1] = ry / rl;
    r[2] = 0.0f;
    u[0] = r[1] * f[2] - r[2] * f[1];
    u[1] = r[2] * f[0] - r[0] * f[2];
    u[2] = r[0] * f[1] - r[1] * f[0];
}

// Camera.swift:5-13.
OrbitCamera camera_from(const Options& o) {
    OrbitCamera c;
    if (o.cam_distance) {
        c.azimuth = (float)*o.cam_azimuth;
        c.elevation = std::fmin(std::fmax((float)*o.cam_elevation, -1.5f), 1.5f);
        c.distance = std::fmin(std::fmax((float)*o.cam_distance, OrbitCamera::kMinimumDistance),
                               OrbitCamera::kMaximumDistance);
    }
    return c;
}

// Schwarzschild.Settings.interactive / .still (Schwarzschild.swift:167-172).
garg_geodesic_settings policy_settings(Policy p) {
    garg_geodesic_settings s;
    s.step_factor = GARG_F(STEP_RADIUS_FACTOR);
    s.step_min = GARG_F(STEP_MIN);
    s.step_max = GARG_F(STEP_MAX);
    s.capture_radius = GARG_F(R_CAPTURE);
    s.escape_radius = GARG_F(R_ESCAPE);
    s.step_cap = p == Policy::still ? STEP_CAP_STILL : STEP_CAP_INTERACTIVE;
    s.empty_step_factor = p == Policy::still ? GARG_F(STEP_FACTOR_EMPTY_STILL) : GARG_F(STEP_FACTOR_EMPTY_INTERACTIVE);
    s.empty_step_max = p == Policy::still ? GARG_F(STEP_MAX_EMPTY_STILL) : GARG_F(STEP_MAX_EMPTY_INTERACTIVE);
    return s;
}

// Probes.cameraGrid (Validation/Probes.swift:52-66): the camera's rays
// through the centres of a columns x rows grid over the field of view.
std::vector<garg_ray_probe> camera_grid(const OrbitCamera& cam, int columns, int rows, double aspect) {
    float pos[3], right[3], up[3], fwd[3];
    cam.basis(pos, right, up, fwd);
    double tan_half = std::tan(cam.fov_y * 0.5f);
    std::vector<garg_ray_probe> out;
    for (int row = 0; row < rows; ++row)
        for (int col = 0; col < columns; ++col) {
            double x = (2.0 * (col + 0.5) / columns - 1.0) * tan_half * aspect;
            double y = (1.0 - 2.0 * (row + 0.5) / rows) * tan_half;
            double d[3], len = 0;
            for (int k = 0; k < 3; ++k) {
                d[k] = fwd[k] + x * right[k] + y * up[k];
                len += d[k] * d[k];
            }
            garg_ray_probe p;
            for (int k = 0; k < 3; ++k) {
                p.origin[k] = pos[k];
                p.direction[k] = (float)(d[k] / std::sqrt(len));
            }
            out.push_back(p);
        }
    return out;
}

// MarchPass.uniforms (MarchPass.swift:67-98) without the table fields.
garg_march_uniforms make_uniforms(const OrbitCamera& camera, int width, int height, Policy policy, bool redshift,
                                  uint32_t star_seed, bool stars, double spin, bool compensated) {
    garg_march_uniforms u;
    memset(&u, 0, sizeof u);
    camera.basis(u.camera_position, u.camera_right, u.camera_up, u.camera_forward);
    float tan_half = std::tan(camera.fov_y * 0.5f);
    u.tan_half_fov[0] = tan_half * (float)width / (float)height;
    u.tan_half_fov[1] = tan_half;
    u.resolution[0] = (uint32_t)width;
    u.resolution[1] = (uint32_t)height;
    u.volume_half_extent = GARG_F(VOLUME_HALF_EXTENT);
    u.opacity_scale = GARG_F(VOLUME_OPACITY);
    u.star_brightness = stars ? GARG_F(STAR_BRIGHTNESS) : 0.0f;
    u.drift_budget = policy == Policy::still ? GARG_F(DRIFT_BUDGET_STILL) : GARG_F(DRIFT_BUDGET_INTERACTIVE);
    u.star_seed = star_seed;
    u.redshift = redshift ? 1u : 0u;
    u.bake_capacity = BAKE_SAMPLES;
    u.spin = (float)spin;
    u.compensated = compensated ? 1u : 0u;
    u.geodesic = policy_settings(policy);
    return u;
}

float half_to_float(uint16_t h) {
    uint32_t sign = (uint32_t)(h >> 15) << 31, exp = (h >> 10) & 0x1Fu, mant = h & 0x3FFu, bits;
    if (exp == 0) {
        if (mant == 0) bits = sign;
        else {  // subnormal: normalize
            int e = -1;
            do {
                ++e;
                mant <<= 1;
            } while (!(mant & 0x400u));
            bits = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3FFu) << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7F800000u | (mant << 13);
    } else {
        bits = sign | ((exp + 112u) << 23) | (mant << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

// ---- PNG --------------------------------------------------------------------

namespace {

uint32_t crc_table[256];

void crc_init() {
    for (uint32_t n = 0; n < 256; ++n) {
        uint32_t c = n;
        for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        crc_table[n] = c;
    }
}

uint32_t crc32(uint32_t crc, const uint8_t* p, size_t n) {
    crc ^= 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) crc = crc_table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return crc ^ 0xFFFFFFFFu;
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put32(out, (uint32_t)data.size());
    size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put32(out, crc32(0, out.data() + start, out.size() - start));
}

}  // namespace

// PNG (ISO/IEC 15948) with zlib (RFC 1950) of stored deflate blocks (RFC 1951).
bool write_png(const std::string& path, const uint8_t* rgba, int width, int height) {
    crc_init();
    std::vector<uint8_t> raw;
    raw.reserve((size_t)height * (width * 4 + 1));
    for (int y = 0; y < height; ++y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), rgba + (size_t)y * width * 4, rgba + (size_t)(y + 1) * width * 4);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    size_t pos = 0;
    do {
        size_t n = std::min<size_t>(65535, raw.size() - pos);
        bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)n);
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)~n);
        z.push_back((uint8_t)(~n >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
        pos += n;
    } while (pos < raw.size());
    put32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)width);
    put32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = fwrite(out.data(), 1, out.size(), f) == out.size();
    return fclose(f) == 0 && ok;
}

// ---- scene --------------------------------------------------------------------

void Scene::close() {
    garg_tables_destroy(tables);
    garg_volume_destroy(volume);
    garg_particles_destroy(particles);
    garg_ctx_destroy(ctx);
    *this = Scene{};
}

// Validate.advance and the splat of Still.render (Still.swift:31-39).
// The exact splat's LDS brick the autotuner chose for this deposit, from
// <autotune_dir>/<arch>/splat_<N>_<size>.json ({"config": {"brick": [x, y, z]}},
// N in millions when whole); false when there is no such entry.
bool tuned_brick(const char* arch, int particles, int volume, uint32_t shape[3]) {
    const std::string count = particles % 1000000 == 0 ? std::to_string(particles / 1000000) + "M"
                                                       : std::to_string(particles);
    const std::string path = autotune_dir() + "/" + arch + "/splat_" + count + "_" + std::to_string(volume) + ".json";
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return false;
    std::string text;
    char buf[512];
    for (size_t n; (n = fread(buf, 1, sizeof buf, f)) > 0;) text.append(buf, n);
    fclose(f);
    size_t at = text.find("\"brick\"");
    if (at == std::string::npos || (at = text.find('[', at)) == std::string::npos) return false;
    const char* p = text.c_str() + at + 1;
    for (int k = 0; k < 3; ++k) {
        char* end = nullptr;
        long v = strtol(p, &end, 10);
        if (end == p || v <= 0 || v > 64) return false;
        shape[k] = (uint32_t)v;
        p = end;
        while (*p == ' ' || *p == ',' || *p == '\n') ++p;
    }
    return true;
}

Scene make_scene(const Configuration& c, int settle_substeps, bool exact_splat) {
    Scene s;
    check(garg_ctx_create(0, 0, &s.ctx), "creating the GPU context");
    garg_device_info info;
    check(garg_device_info_get(s.ctx, &info), "reading the device");
    std::string refusal = budget_refusal(c.bytes(), info.budget_bytes,
                                         std::string("Preset ") + preset_name(c.preset));
    if (!refusal.empty()) operational(refusal);
    check(garg_particles_create(s.ctx, (uint32_t)c.particles, c.seed, GARG_POTENTIAL_PACZYNSKI_WIITA, &s.particles),
          "allocating particles");
    check(garg_particles_seed(s.particles, nullptr, GARG_GENERATOR_DISK, 0.0f), "seeding");
    check(garg_particles_step(s.particles, nullptr, (uint32_t)(settle_substeps / (int)SIM_SUBSTEPS)), "settling");
    check(garg_volume_create(s.ctx, (uint32_t)c.volume, GARG_F(VOLUME_HALF_EXTENT), 0.0f, &s.volume),
          "allocating the volume");
    uint32_t brick[3];
    if (exact_splat && tuned_brick(info.arch, c.particles, c.volume, brick))
        check(garg_volume_set_splat_brick(s.volume, brick[0], brick[1], brick[2]), "setting the tuned brick");
    check(garg_splat_bin(s.volume, s.particles, nullptr, exact_splat ? 1u : 0u, 0), "splatting");
    check(garg_splat_resolve(s.volume, nullptr, 1u), "resolving");
    check(garg_tables_create(s.ctx, nullptr, GARG_F(VOLUME_HALF_EXTENT), &s.tables), "building sweep tables");
    check(garg_stream_sync(s.ctx, nullptr), "settling the scene");
    return s;
}

// Where the batch autotuner's table lives: GARG_AUTOTUNE_DIR, else the
// autotune directory beside libgargantua_hip.so.
std::string autotune_dir() {
    if (const char* env = getenv("GARG_AUTOTUNE_DIR"); env && *env) return env;
    Dl_info info;
    if (dladdr((void*)&garg_version, &info) && info.dli_fname) {
        std::string lib = info.dli_fname;
        size_t slash = lib.find_last_of('/');
        return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/autotune";
    }
    return "autotune";
}

}  // namespace host
