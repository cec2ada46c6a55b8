// `gargantua validate` (Validation/Validate.swift:19-387): runs the goldens
// of goldens/*.json on this GPU and prints a table; exit 0 when every judged
// golden passes, 1 when one fails, 2 on an operational error.
//
// The product never links the FP64 CPU oracle. Parity and sky, whose bar
// is the reference's FP64 integrator, are judged against the library's FP64
// GPU run of that integrator (garg_probe_fp64), and beaming, whose bar is
// the Oracle's FP64 lensed render with the gas, against the library's FP64
// GPU render (garg_render_fp64). The other goldens are judged by closed
// forms (shadow, deflection), budgets (conservation), bitwise equality
// (determinism), the disk edge fit (isco), GPU against GPU (jump), and
// shadow_kerr against the GPU's FP64 bisection.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <unistd.h>
#include <vector>

#include "../../include/gargantua_constants.h"
#include "host.hpp"

namespace host {

namespace {

struct Golden {
    std::string name, claim, kind, method;
    double expected = 0, tolerance = 0;
    std::map<std::string, double> parameters;
    bool passes(double m) const {
        if (kind == "relative") return std::fabs(m - expected) / std::fabs(expected) <= tolerance;
        if (kind == "absolute") return std::fabs(m - expected) <= tolerance;
        return m == expected;
    }
    double p(const char* key) const {
        auto it = parameters.find(key);
        if (it == parameters.end()) operational("golden " + name + " has no parameter " + key + ".");
        return it->second;
    }
};

// Reader for the flat golden format (Oracle/Golden.swift:5-51): string and
// number members plus one object of numbers.
struct Reader {
    const std::string& s;
    size_t i = 0;
    void ws() {
        while (i < s.size() && isspace((unsigned char)s[i])) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    std::string str() {
        ws();
        std::string out;
        if (s[i] != '"') return out;
        for (++i; i < s.size() && s[i] != '"'; ++i) {
            if (s[i] == '\\' && i + 1 < s.size()) ++i;
            out += s[i];
        }
        ++i;
        return out;
    }
    double num() {
        ws();
        char* end = nullptr;
        double v = strtod(s.c_str() + i, &end);
        i = end - s.c_str();
        return v;
    }
};

Golden load(const std::string& dir, const std::string& name) {
    std::string path = dir + "/" + name + ".json";
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) operational("Golden " + name + " is missing from " + dir + ".");
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    fclose(f);
    Golden g;
    Reader r{text};
    if (!r.eat('{')) operational("Golden " + name + " does not parse.");
    do {
        std::string key = r.str();
        if (!r.eat(':')) operational("Golden " + name + " does not parse.");
        if (key == "parameters") {
            r.eat('{');
            if (!r.eat('}')) {
                do {
                    std::string k = r.str();
                    r.eat(':');
                    g.parameters[k] = r.num();
                } while (r.eat(','));
                r.eat('}');
            }
        } else if (key == "expected") g.expected = r.num();
        else if (key == "tolerance") g.tolerance = r.num();
        else {
            std::string v = r.str();
            if (key == "name") g.name = v;
            else if (key == "claim") g.claim = v;
            else if (key == "kind") g.kind = v;
            else if (key == "method") g.method = v;
        }
    } while (r.eat(','));
    return g;
}

std::string goldens_dir() {
    char exe[4096];
    ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 1);
    std::string base = n > 0 ? std::string(exe, n) : std::string(".");
    base = base.substr(0, base.find_last_of('/'));
    for (std::string c : {base + "/../../tests/golden", base + "/../tests/golden", std::string("tests/golden")})
        if (access((c + "/isco.json").c_str(), R_OK) == 0) return c;
    operational("The goldens directory (tests/golden) was not found.");
}

struct Row {
    Golden g;
    std::string measured;
    std::string status;  // pass, FAIL, oracle
};

// Schwarzschild launch from radius R toward the hole with impact parameter b
// (Schwarzschild.swift:100-104): origin (R, 0, 0), static-observer direction.
garg_ray_probe aimed(double radius, double b) {
    double dphi = b * std::sqrt(1.0 - 2.0 / radius) / radius;
    double dr = -std::sqrt(std::max(0.0, 1.0 - dphi * dphi));
    garg_ray_probe p;
    p.origin[0] = (float)radius;
    p.origin[1] = p.origin[2] = 0.0f;
    p.direction[0] = (float)dr;
    p.direction[1] = (float)dphi;
    p.direction[2] = 0.0f;
    return p;
}

std::vector<garg_ray_probe_result> probe(garg_ctx* ctx, const std::vector<garg_ray_probe>& in,
                                         const garg_geodesic_settings& s, double spin, garg_tables* tables) {
    OrbitCamera cam;
    garg_march_uniforms u = make_uniforms(cam, 1, 1, Policy::still, false, 0, false, spin, true);
    u.geodesic = s;
    u.drift_budget = 1.0f;
    if (tables) check(garg_tables_apply(tables, &u), "applying tables");
    std::vector<garg_ray_probe_result> out(in.size());
    check(garg_probe(ctx, nullptr, in.data(), (uint32_t)in.size(), &u, tables, GARG_INTEGRATOR_FAST, out.data()),
          "probing");
    return out;
}

double inner_edge(const std::vector<double>& radii, double bw, double lo, double hi) {
    int count = (int)std::lround((hi - lo) / bw);
    std::vector<double> counts(count, 0.0);
    for (double r : radii)
        if (r >= lo && r < hi) counts[std::min((int)((r - lo) / bw), count - 1)] += 1.0;
    double sx = 0, sy = 0, sxx = 0, sxy = 0;
    for (int k = 0; k < count; ++k) {
        double c = lo + (k + 0.5) * bw;
        double y = counts[k] / (2.0 * M_PI * c * bw) * std::pow(c, 1.5);
        sx += c;
        sy += y;
        sxx += c * c;
        sxy += c * y;
    }
    double slope = (count * sxy - sx * sy) / (count * sxx - sx * sx);
    return -((sy - slope * sx) / count) / slope;
}

std::string fmt(const char* f, double a, double b = 0, double c = 0) {
    char buf[256];
    snprintf(buf, sizeof buf, f, a, b, c);
    return buf;
}

}  // namespace

void run_validate(const Options& o) {
    const std::string dir = goldens_dir();
    Configuration c = Configuration::make(o);
    Golden isco_g = load(dir, "isco");
    Scene s = make_scene(c, (int)isco_g.p("settlingSteps"), true);
    std::vector<Row> rows;
    garg_device_info info;
    check(garg_device_info_get(s.ctx, &info), "reading the device");

    // isco (Validate.swift:102-110).
    {
        std::vector<garg_particle> ps(c.particles);
        check(garg_particles_read(s.particles, ps.data(), (uint32_t)ps.size()), "reading particles");
        std::vector<double> radii;
        for (auto& p : ps) radii.push_back(std::hypot((double)p.position[0], (double)p.position[1]));
        double edge = inner_edge(radii, isco_g.p("binWidth"), isco_g.p("fitInner"), isco_g.p("fitOuter"));
        rows.push_back({isco_g, fmt("r = %.3f", edge), isco_g.passes(edge) ? "pass" : "FAIL"});
    }
    // determinism (Validate.swift:112-122).
    {
        Golden g = load(dir, "determinism");
        uint64_t sums[2];
        for (int k = 0; k < 2; ++k) {
            garg_particles* p = nullptr;
            check(garg_particles_create(s.ctx, (uint32_t)c.particles, c.seed, GARG_POTENTIAL_PACZYNSKI_WIITA, &p), "alloc");
            check(garg_particles_seed(p, nullptr, GARG_GENERATOR_DISK, 0.0f), "seeding");
            check(garg_particles_step(p, nullptr, (uint32_t)(g.p("steps") / 2)), "stepping");
            check(garg_particles_checksum(p, &sums[k]), "checksum");
            garg_particles_destroy(p);
        }
        char buf[64];
        snprintf(buf, sizeof buf, "%016llx vs %016llx", (unsigned long long)sums[0], (unsigned long long)sums[1]);
        rows.push_back({g, buf, sums[0] == sums[1] ? "pass" : "FAIL"});
    }
    // shadow (Validate.swift:125-150).
    {
        Golden g = load(dir, "shadow");
        double low = g.p("bracketLow"), high = g.p("bracketHigh"), radius = g.p("launchRadius");
        int n = (int)g.p("raysPerRound");
        double measured = -1;
        for (int round = 0; round < (int)g.p("rounds") && measured < 0; ++round) {
            std::vector<double> values;
            std::vector<garg_ray_probe> in;
            for (int i = 0; i < n; ++i) {
                values.push_back(low + (high - low) * (i + 0.5) / n);
                in.push_back(aimed(radius, values.back()));
            }
            auto res = probe(s.ctx, in, policy_settings(Policy::still), 0.0, nullptr);
            int first = n;
            for (int i = 0; i < n; ++i) {
                if (res[i].outcome == GARG_RAY_EXHAUSTED) {
                    measured = values[i];
                    break;
                }
                if (res[i].outcome == GARG_RAY_ESCAPED && first == n) first = i;
            }
            if (measured >= 0) break;
            double nl = first > 0 ? values[first - 1] : low, nh = first < n ? values[first] : high;
            low = nl;
            high = nh;
        }
        double b = measured >= 0 ? measured : 0.5 * (low + high);
        rows.push_back({g, fmt("b = %.4f", b), g.passes(b) ? "pass" : "FAIL"});
    }
    // deflection (Validate.swift:152-167).
    {
        Golden g = load(dir, "deflection");
        double radius = g.p("launchRadius"), b = g.p("impactParameter");
        garg_geodesic_settings st = policy_settings(Policy::still);
        st.step_max = 1e9f;
        st.escape_radius = (float)radius;
        garg_ray_probe launch = aimed(radius, b);
        auto r = probe(s.ctx, {launch}, st, 0.0, nullptr)[0];
        double dphi = b * std::sqrt(1.0 - 2.0 / radius) / radius, dr = -std::sqrt(1.0 - dphi * dphi);
        double a[3] = {dr, dphi, 0.0}, f[3] = {r.direction[0], r.direction[1], r.direction[2]};
        double cx = a[1] * f[2] - a[2] * f[1], cy = a[2] * f[0] - a[0] * f[2], cz = a[0] * f[1] - a[1] * f[0];
        double angle = std::atan2(std::sqrt(cx * cx + cy * cy + cz * cz), a[0] * f[0] + a[1] * f[1] + a[2] * f[2]);
        bool ok = g.passes(angle) && r.outcome == GARG_RAY_ESCAPED;
        rows.push_back({g, fmt("%.6f rad", angle), ok ? "pass" : "FAIL"});
    }
    // conservation (Validate.swift:169-177).
    for (const char* kind : {"interactive", "still"}) {
        Golden g = load(dir, std::string("conservation_") + kind);
        auto launches = camera_grid(OrbitCamera{}, (int)g.p("columns"), (int)g.p("rows"), 960.0 / 540.0);
        auto res = probe(s.ctx, launches, policy_settings(kind[0] == 'i' ? Policy::interactive : Policy::still), 0.0,
                         nullptr);
        double worst = 0;
        int exhausted = 0;
        for (auto& r : res) {
            worst = std::max(worst, (double)r.drift);
            exhausted += r.outcome == GARG_RAY_EXHAUSTED;
        }
        char buf[128];
        snprintf(buf, sizeof buf, "max drift %.2e, %d of %zu rays exhausted", worst, exhausted, res.size());
        rows.push_back({g, buf, g.passes(worst) ? "pass" : "FAIL"});
    }
    // parity (Validate.swift:179-200): the GPU's probe end points against the
    // reference's FP64 integrator, here the library's FP64 GPU probe, from the
    // same float launches.
    {
        Golden g = load(dir, "parity");
        auto launches = camera_grid(OrbitCamera{}, (int)g.p("columns"), (int)g.p("rows"), 960.0 / 540.0);
        const garg_geodesic_settings still = policy_settings(Policy::still);
        auto res = probe(s.ctx, launches, still, 0.0, nullptr);
        std::vector<garg_ray_probe_fp64_result> ref(launches.size());
        check(garg_probe_fp64(s.ctx, nullptr, launches.data(), (uint32_t)launches.size(), &still, ref.data()),
              "integrating the FP64 probes");
        double worst = 0;
        int mismatched = 0;
        for (size_t k = 0; k < res.size(); ++k) {
            double d2 = 0, e2 = 0;
            for (int c = 0; c < 3; ++c) {
                const double e = ref[k].position[c], a = res[k].position[c];
                d2 += (e - a) * (e - a);
                e2 += e * e;
            }
            worst = std::max(worst, std::sqrt(d2) / std::max(std::sqrt(e2), 1.0));
            mismatched += res[k].outcome != ref[k].outcome;
        }
        rows.push_back({g, fmt("max relative endpoint error %.2e, %.0f outcome mismatches", worst, mismatched),
                        g.passes(worst) && mismatched == 0 ? "pass" : "FAIL"});
    }
    // beaming (Validate.swift:258-291): the brightness ratio of the image
    // halves of the GPU's still march against that of the Oracle's FP64
    // lensed render of the same resolved volume, here the library's FP64 GPU
    // render (garg_render_fp64).
    {
        Golden g = load(dir, "beaming");
        const int w = (int)g.p("width"), h = (int)g.p("height");
        OrbitCamera cam;
        garg_march_uniforms u = make_uniforms(cam, w, h, Policy::still, true, 0, false, 0.0, true);
        garg_targets* t = nullptr;
        check(garg_targets_create(s.ctx, w, h, &t), "allocating targets");
        check(garg_march_image(s.ctx, nullptr, &u, s.volume, nullptr, t, 0, 0, GARG_INTEGRATOR_REFERENCE), "marching");
        std::vector<uint16_t> color((size_t)w * h * 4);
        check(garg_targets_read(t, color.data(), nullptr, nullptr), "reading");
        garg_targets_destroy(t);
        std::vector<double> lum((size_t)w * h), oracle_lum((size_t)w * h);
        for (size_t i = 0; i < lum.size(); ++i) {
            double rgb[3];
            for (int k = 0; k < 3; ++k) rgb[k] = half_to_float(color[4 * i + k]);
            lum[i] = 0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2];
        }
        garg_render_camera_fp64 oc{};
        float position[3], right[3], up[3], forward[3];
        cam.basis(position, right, up, forward);
        for (int k = 0; k < 3; ++k) {
            oc.position[k] = position[k];
            oc.right[k] = right[k];
            oc.up[k] = up[k];
            oc.forward[k] = forward[k];
        }
        const double tan_half = std::tan((double)cam.fov_y * 0.5);
        oc.tan_half_fov[0] = tan_half * w / h;
        oc.tan_half_fov[1] = tan_half;
        const garg_geodesic_settings still = policy_settings(Policy::still);
        std::vector<double> rgb((size_t)w * h * 3);
        check(garg_render_fp64(s.ctx, nullptr, s.volume, &oc, (uint32_t)w, (uint32_t)h, &still, VOLUME_OPACITY, 1u,
                               rgb.data()),
              "rendering in FP64");
        for (size_t i = 0; i < oracle_lum.size(); ++i)
            oracle_lum[i] = 0.2126 * rgb[3 * i] + 0.7152 * rgb[3 * i + 1] + 0.0722 * rgb[3 * i + 2];
        // LensedRender.sideRatio (LensedRender.swift:99-109): the brighter
        // half over the dimmer, split at the column through the hole.
        auto side_ratio = [&](const std::vector<double>& l) {
            double left = 0, right_sum = 0;
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x) (x < w / 2 ? left : right_sum) += l[(size_t)y * w + x];
            const double high = std::max(left, right_sum), low = std::min(left, right_sum);
            return low > 0.0 ? high / low : INFINITY;
        };
        const double gpu = side_ratio(lum), fp64 = side_ratio(oracle_lum), m = gpu / fp64;
        rows.push_back({g, fmt("gpu %.3f vs fp64 %.3f, ratio %.3f", gpu, fp64, m), g.passes(m) ? "pass" : "FAIL"});
    }
    // sky (Validate.swift:293-320): the sweep tables' asymptotic directions
    // against the FP64 integrator run out to oracleRadius.
    {
        Golden g = load(dir, "sky");
        OrbitCamera cam;
        check(garg_tables_update(s.tables, nullptr, cam.distance), "updating tables");
        auto launches = camera_grid(cam, (int)g.p("columns"), (int)g.p("rows"), 960.0 / 540.0);
        auto res = probe(s.ctx, launches, policy_settings(Policy::still), 0.0, s.tables);
        garg_geodesic_settings far = policy_settings(Policy::still);
        far.step_factor = 0.01f;
        far.step_max = 20.0f;
        far.escape_radius = (float)g.p("oracleRadius");
        far.step_cap = 400000u;
        far.empty_step_factor = (float)STEP_RADIUS_FACTOR;
        far.empty_step_max = 20.0f;
        std::vector<garg_ray_probe_fp64_result> ref(launches.size());
        check(garg_probe_fp64(s.ctx, nullptr, launches.data(), (uint32_t)launches.size(), &far, ref.data()),
              "integrating the FP64 probes");
        double worst = 0;
        int mismatched = 0, compared = 0;
        for (size_t k = 0; k < res.size(); ++k) {
            const bool escaped = res[k].outcome == GARG_RAY_ESCAPED;
            if (escaped != (ref[k].outcome == GARG_RAY_ESCAPED)) {
                ++mismatched;
                continue;
            }
            if (!escaped) continue;
            double dot = 0;
            for (int c = 0; c < 3; ++c) dot += ref[k].direction[c] * res[k].direction[c];
            worst = std::max(worst, std::acos(std::min(std::max(dot, -1.0), 1.0)));
            ++compared;
        }
        rows.push_back({g,
                        fmt("max sky angle error %.2e rad over %.0f escaping rays, %.0f fate mismatches", worst, compared,
                            mismatched),
                        g.passes(worst) && mismatched == 0 && compared > 0 ? "pass" : "FAIL"});
    }
    // jump (Validate.swift:323-346).
    {
        Golden g = load(dir, "jump");
        int w = (int)g.p("width"), h = (int)g.p("height");
        OrbitCamera cam;
        check(garg_tables_update(s.tables, nullptr, cam.distance), "updating tables");
        std::vector<double> lum[2];
        for (int jump = 0; jump < 2; ++jump) {
            garg_march_uniforms u = make_uniforms(cam, w, h, Policy::still, true, 0, false, 0.0, true);
            if (jump) check(garg_tables_apply(s.tables, &u), "applying tables");
            garg_targets* t = nullptr;
            check(garg_targets_create(s.ctx, w, h, &t), "allocating targets");
            check(garg_march_image(s.ctx, nullptr, &u, s.volume, jump ? s.tables : nullptr, t, 0, 0,
                                   GARG_INTEGRATOR_FAST),
                  "marching");
            std::vector<uint16_t> color((size_t)w * h * 4);
            check(garg_targets_read(t, color.data(), nullptr, nullptr), "reading");
            garg_targets_destroy(t);
            for (size_t i = 0; i < (size_t)w * h; ++i) {
                float rgb[3];
                for (int k = 0; k < 3; ++k) rgb[k] = half_to_float(color[4 * i + k]);
                lum[jump].push_back(0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2]);
            }
        }
        double diff = 0, total = 0;
        for (size_t i = 0; i < lum[0].size(); ++i) {
            diff += std::fabs(lum[0][i] - lum[1][i]);
            total += lum[0][i];
        }
        double m = total > 0 ? diff / total : 1.0;
        rows.push_back({g, fmt("mean relative difference %.2e", m), g.passes(m) ? "pass" : "FAIL"});
    }
    // shadow_kerr: fp32 probe bisection against the GPU's FP64 bisection.
    {
        Golden g = load(dir, "shadow_kerr");
        double spin = g.p("spin"), radius = g.p("launchRadius");
        int iterations = (int)g.p("iterations");
        double xi[2];
        for (int side = 0; side < 2; ++side) {
            double sign = side == 0 ? -1.0 : 1.0, low = 0.0, high = 12.0 / radius;
            garg_geodesic_settings st = policy_settings(Policy::still);
            for (int it = 0; it < iterations; ++it) {
                double aim = 0.5 * (low + high);
                garg_ray_probe p{{(float)radius, 0, 0}, {-1.0f, (float)(sign * aim), 0}};
                auto r = probe(s.ctx, {p}, st, spin, nullptr)[0];
                (r.outcome == GARG_RAY_CAPTURED ? low : high) = aim;
            }
            // The edge is measured as an aim tan; xi follows from the FP64 kernel's
            // equatorial boundary at the matching screen azimuth below.
            xi[side] = 0.5 * (low + high);
        }
        std::vector<double> edge(360), kxi(360);
        check(garg_shadow_boundary(s.ctx, nullptr, spin, radius, 90.0, 360, 6, 64, STEP_CAP_STILL, edge.data(),
                                   kxi.data()),
              "bisecting in FP64");
        // aim tan -> angle; compare edges (retrograde over prograde) in angle.
        double gpu = std::atan(xi[0]) / std::atan(xi[1]);
        double fp64 = edge[180] / edge[0];
        double m = gpu / fp64;
        char buf[160];
        snprintf(buf, sizeof buf, "fp32 %.4f vs fp64 %.4f, ratio %.4f (xi asymmetry %.4f)", gpu, fp64, m,
                 std::fabs(kxi[180] / kxi[0]));
        rows.push_back({g, buf, g.passes(m) ? "pass" : "FAIL"});
    }

    line(std::string("validation on ") + info.name + " (" + info.arch + ")");
    bool all = true;
    for (auto& r : rows) {
        char buf[512];
        snprintf(buf, sizeof buf, "%-25s %-60s %-48s %-8s", r.g.name.c_str(), r.g.claim.substr(0, 60).c_str(),
                 r.measured.c_str(), r.status.c_str());
        line(buf);
        if (r.status == "FAIL") all = false;
    }
    s.close();
    exit(all ? 0 : 1);
}

}  // namespace host
