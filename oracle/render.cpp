// FP64 oracle: volume field sampling, the lensed render of LensedRender.swift,
// and a per-pixel restatement of the march kernel (marchImage / probeRays)
// including the sphere jump, the gather and the starfield. Test
// infrastructure only (see oracle.h).
#include "oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>

using orc::v3;

extern "C" {

// VolumeField.swift:40-63: trilinear on texel centres, zero outside; the z
// axis may span its own half extent.
void orc_field_sample(const orc_field* field, const double position[3], double e_out[4], double v_out[3]) {
    int n = field->size;
    double half[3] = {field->half_extent, field->half_extent, field->half_extent_z};
    double t[3], frac[3];
    int base[3];
    for (int i = 0; i < 3; ++i) {
        double unit = (position[i] + half[i]) / (2.0 * half[i]);
        t[i] = unit * (double)n - 0.5;
        double b = std::floor(t[i]);
        frac[i] = t[i] - b;
        base[i] = (int)b;
    }
    for (int i = 0; i < 4; ++i) e_out[i] = 0.0;
    for (int i = 0; i < 3; ++i) v_out[i] = 0.0;
    for (int dz = 0; dz <= 1; ++dz) {
        double wz = dz == 0 ? 1.0 - frac[2] : frac[2];
        for (int dy = 0; dy <= 1; ++dy) {
            double wy = dy == 0 ? 1.0 - frac[1] : frac[1];
            for (int dx = 0; dx <= 1; ++dx) {
                double wx = dx == 0 ? 1.0 - frac[0] : frac[0];
                double w = wx * wy * wz;
                if (w == 0.0) continue;
                int x = base[0] + dx, y = base[1] + dy, z = base[2] + dz;
                if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) continue;
                int64_t i = ((int64_t)z * n + y) * n + x;
                for (int c = 0; c < 4; ++c) e_out[c] += field->emission[4 * i + c] * w;
                for (int c = 0; c < 3; ++c) v_out[c] += field->velocity[3 * i + c] * w;
            }
        }
    }
}

// LensedRender.swift:94-109.
double orc_luminance(const double rgb[3]) { return 0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2]; }

double orc_side_ratio(const double* luminance, int width, int height) {
    double left = 0.0, right = 0.0;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            if (x < width / 2) left += luminance[y * width + x];
            else right += luminance[y * width + x];
        }
    double high = std::max(left, right), low = std::min(left, right);
    return low > 0.0 ? high / low : INFINITY;
}

// LensedRender.swift:87-92.
void orc_camera_direction(const orc_camera* c, int px, int py, int width, int height, double out[3]) {
    double x = (2.0 * ((double)px + 0.5) / (double)width - 1.0) * c->tan_half_fov[0];
    double y = (1.0 - 2.0 * ((double)py + 0.5) / (double)height) * c->tan_half_fov[1];
    v3 d = orc::make(c->forward) + x * orc::make(c->right) + y * orc::make(c->up);
    orc::store(orc::normalize(d), out);
}

static double step_length(const orc_settings* g, double r, double z) {
    double coarse = std::min(std::max(g->empty_step_factor * r, g->step_min), g->empty_step_max);
    if (std::fabs(z) > DISK_SLAB_HALF_HEIGHT + coarse) return coarse;
    return std::min(std::max(g->step_factor * r, g->step_min), g->step_max);
}

// LensedRender.swift:46-78.
void orc_trace(const double origin[3], const double direction[3], const orc_field* field, const orc_settings* settings,
               double opacity, double slab, int redshift, double out[3]) {
    orc_ray ray;
    orc_schw_launch(origin, direction, &ray);
    v3 color = {0, 0, 0};
    double transmittance = 1.0;
    double prev[3];
    orc_schw_position(&ray, prev);
    for (int64_t k = 0; k < settings->step_cap; ++k) {
        orc_schw_step(&ray.s, step_length(settings, ray.s.r, prev[2]));
        double pos[3];
        orc_schw_position(&ray, pos);
        if (ray.s.r <= settings->capture_radius) break;
        if (std::fabs(pos[2]) < slab && std::fabs(pos[0]) < field->half_extent && std::fabs(pos[1]) < field->half_extent &&
            transmittance > TRANSMITTANCE_FLOOR) {
            double e[4], v[3];
            orc_field_sample(field, pos, e, v);
            if (e[3] > 0.0) {
                v3 dp = orc::make(pos) - orc::make(prev);
                double path = orc::length(dp);
                double g3 = 1.0;
                if (redshift) {
                    double d[3];
                    orc_schw_direction(&ray, d);
                    double f = orc_schw_f(ray.s.r);
                    double g = orc_redshift_factor(f, d, v);
                    double gravity = std::sqrt(std::max(f, 0.0));
                    double bulk = std::min(e[3] / BEAMING_DENSITY_FLOOR, 1.0);
                    g3 = gravity * gravity * gravity * (1.0 - bulk) + g * g * g * bulk;
                }
                double alpha = 1.0 - std::exp(-opacity * e[3] * path);
                color = color + (transmittance * g3 * path) * v3{e[0], e[1], e[2]};
                transmittance *= 1.0 - alpha;
            }
        }
        prev[0] = pos[0];
        prev[1] = pos[1];
        prev[2] = pos[2];
        if (ray.s.r >= settings->escape_radius) break;
    }
    orc::store(color, out);
}

// LensedRender.swift:81-91, multi-threaded over rows.
void orc_render(int width, int height, const orc_camera* camera, const orc_field* field, const orc_settings* settings,
                double opacity, int redshift, int threads, double* out_rgb) {
    orc::parallel_dynamic(height, threads, 1, [&](int64_t y0, int64_t y1) {
        for (int64_t y = y0; y < y1; ++y)
            for (int x = 0; x < width; ++x) {
                double d[3];
                orc_camera_direction(camera, x, (int)y, width, height, d);
                orc_trace(camera->position, d, field, settings, opacity, DISK_SLAB_HALF_HEIGHT, redshift,
                          out_rgb + 3 * ((int64_t)y * width + x));
            }
    });
}

// ---- march kernel restatement --------------------------------------------

// Linear sampling of an N entry table at a normalized coordinate with texel
// centres at (i + 0.5) / N and clamp to edge (how the kernel samples its
// sweep and blackbody tables).
static double table_sample(const double* table, int n, int stride, int channel, double coordinate) {
    double x = coordinate * (double)n - 0.5;
    double b = std::floor(x);
    double f = x - b;
    int i0 = (int)b, i1 = i0 + 1;
    i0 = std::min(std::max(i0, 0), n - 1);
    i1 = std::min(std::max(i1, 0), n - 1);
    return table[i0 * stride + channel] * (1.0 - f) + table[i1 * stride + channel] * f;
}

// Geodesic.h:106-110.
static double sweep_coordinate(double low, double high, double b) {
    double x = std::min(std::max((b - low) / (high - low), 0.0), 1.0);
    double t = 1.0 - std::sqrt(1.0 - x);
    return (t * (SWEEP_TABLE_SIZE - 1.0) + 0.5) / SWEEP_TABLE_SIZE;
}

static double sweep_lookup(const double* table, double low, double high, double b) {
    return table_sample(table, (int)SWEEP_TABLE_SIZE, 1, 0, sweep_coordinate(low, high, b));
}

static v3 asymptotic(const orc_ray& ray, double phi) {
    return std::cos(phi) * orc::make(ray.e1) + std::sin(phi) * orc::make(ray.e2);
}

// Geodesic.h:190-215.
void orc_starfield(const double direction[3], uint32_t seed, const double* blackbody, double out[3]) {
    v3 d = orc::make(direction);
    v3 a = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
    uint32_t face;
    double st[2], major;
    if (a.x >= a.y && a.x >= a.z) {
        face = d.x > 0 ? 0u : 1u;
        st[0] = d.y;
        st[1] = d.z;
        major = a.x;
    } else if (a.y >= a.z) {
        face = d.y > 0 ? 2u : 3u;
        st[0] = d.x;
        st[1] = d.z;
        major = a.y;
    } else {
        face = d.z > 0 ? 4u : 5u;
        st[0] = d.x;
        st[1] = d.y;
        major = a.z;
    }
    double scaled[2];
    uint32_t cell[2];
    for (int i = 0; i < 2; ++i) {
        scaled[i] = (st[i] / major * 0.5 + 0.5) * STAR_CELLS;
        cell[i] = (uint32_t)std::min(scaled[i], STAR_CELLS - 1.0);
    }
    uint32_t in[4] = {cell[0], cell[1], face, seed}, h[4];
    orc_pcg4d(in, h);
    out[0] = out[1] = out[2] = 0.0;
    if (h[0] > STAR_THRESHOLD) return;
    double dx = scaled[0] - ((double)cell[0] + orc_unit(h[1]));
    double dy = scaled[1] - ((double)cell[1] + orc_unit(h[2]));
    double falloff = std::exp(-(dx * dx + dy * dy) * 40.0);
    double brightness = std::pow(orc_unit(h[3]), 3.0) * falloff;
    double temperature = 3000.0 + (12000.0 - 3000.0) * orc_unit(h[3] ^ h[1]);
    double coordinate = orc_blackbody_coordinate(temperature);
    for (int c = 0; c < 3; ++c)
        out[c] = table_sample(blackbody, (int)BLACKBODY_TABLE_SIZE, 3, c, coordinate) * brightness;
}

struct Gather {
    v3 color = {0, 0, 0};
    double transmittance = 1.0, travelled = 0.0, weighted_depth = 0.0, depth_weight = 0.0;

    // march.metal:81-102 (and 151-170 for Kerr).
    void add(const orc_march_params* p, const orc_field* field, const double pos[3], double path, double f,
             const double direction[3]) {
        if (!field) return;
        if (!(std::fabs(pos[2]) < DISK_SLAB_HALF_HEIGHT && std::fabs(pos[0]) < p->volume_half_extent &&
              std::fabs(pos[1]) < p->volume_half_extent && transmittance > TRANSMITTANCE_FLOOR))
            return;
        double e[4], v[3];
        orc_field_sample(field, pos, e, v);
        if (!(e[3] > 0.0)) return;
        double g3 = 1.0;
        if (p->redshift) {
            double g = orc_redshift_factor(f, direction, v);
            double gravity = std::sqrt(std::max(f, 0.0));
            double bulk = std::min(e[3] / BEAMING_DENSITY_FLOOR, 1.0);
            g3 = gravity * gravity * gravity + (g * g * g - gravity * gravity * gravity) * bulk;
        }
        double alpha = 1.0 - std::exp(-p->opacity_scale * e[3] * path);
        v3 contribution = (transmittance * g3 * path) * v3{e[0], e[1], e[2]};
        color = color + contribution;
        double weight = 0.2126 * contribution.x + 0.7152 * contribution.y + 0.0722 * contribution.z;
        weighted_depth += weight * travelled;
        depth_weight += weight;
        transmittance *= 1.0 - alpha;
    }
};

struct SchwMarch {
    orc_ray ray;
    bool sky_only = false;
    v3 sky = {0, 0, 0};
};

// Geodesic.h:130-163: jump from outside the sphere using the sweep tables.
static SchwMarch prepare(const orc_march_params* p, const double origin[3], const double direction[3]) {
    SchwMarch m;
    orc_schw_launch(origin, direction, &m.ray);
    if (p->sphere_radius <= 0.0 || m.ray.s.r <= p->sphere_radius) return m;
    double b = std::fabs(m.ray.angular_momentum) / m.ray.energy;
    if (m.ray.s.rdot >= 0.0) {
        double sweep = sweep_lookup(p->camera_table, 0.0, p->camera_max_b, b);
        m.sky_only = true;
        m.sky = asymptotic(m.ray, m.ray.s.phi + sweep);
        return m;
    }
    if (b >= p->sphere_max_b) {
        double sweep = sweep_lookup(p->sky_table, p->sky_min_b, p->sky_max_b, b);
        m.sky_only = true;
        m.sky = asymptotic(m.ray, m.ray.s.phi + sweep);
        return m;
    }
    double sweep = sweep_lookup(p->sphere_table, 0.0, p->sphere_table_max_b, b) -
                   sweep_lookup(p->camera_table, 0.0, p->camera_max_b, b);
    double r = p->sphere_radius, f = orc_schw_f(r);
    double E = m.ray.energy, L = m.ray.angular_momentum;
    double phidot = L / (r * r);
    double rdot = -std::sqrt(std::max(E * E - f * L * L / (r * r), 0.0));
    double vdot = (E + rdot) / f;
    m.ray.s = {0.0, r, m.ray.s.phi + sweep, vdot, rdot, phidot};
    return m;
}

static double schw_drift(const orc_ray& ray) {
    double e_drift = std::fabs(orc_schw_energy(&ray.s) - ray.energy) / std::fabs(ray.energy);
    double l_drift = std::fabs(orc_schw_angular_momentum(&ray.s) - ray.angular_momentum) /
                     std::max(std::fabs(ray.angular_momentum), 1e-3);
    return std::max(e_drift, l_drift);
}

// march.metal:51-117: the shared loop with sphere landing.
static void march_schw(const orc_march_params* p, const orc_field* field, orc_ray& ray, Gather& gather,
                       int& outcome, int64_t& steps, bool& left_sphere) {
    double prev[3];
    orc_schw_position(&ray, prev);
    outcome = 2;
    left_sphere = false;
    steps = 0;
    const orc_settings* g = &p->geodesic;
    for (int64_t k = 0; k < g->step_cap; ++k) {
        double h = step_length(g, ray.s.r, prev[2]);
        orc_ray_state before = ray.s;
        orc_schw_step(&ray.s, h);
        steps = k + 1;
        if (p->sphere_radius > 0.0 && ray.s.r >= p->sphere_radius && ray.s.rdot > 0.0 && before.r < p->sphere_radius) {
            double frac = (p->sphere_radius - before.r) / (ray.s.r - before.r);
            ray.s = before;
            orc_schw_step(&ray.s, h * frac);
        }
        double pos[3];
        orc_schw_position(&ray, pos);
        double path = orc::length(orc::make(pos) - orc::make(prev));
        gather.travelled += path;
        if (ray.s.r <= g->capture_radius) {
            outcome = 0;
            break;
        }
        double d[3];
        orc_schw_direction(&ray, d);
        gather.add(p, field, pos, path, orc_schw_f(ray.s.r), d);
        prev[0] = pos[0];
        prev[1] = pos[1];
        prev[2] = pos[2];
        if (p->sphere_radius > 0.0 && ray.s.r >= p->sphere_radius && ray.s.rdot > 0.0) {
            outcome = 1;
            left_sphere = true;
            break;
        }
        if (ray.s.r >= g->escape_radius) {
            outcome = 1;
            break;
        }
    }
}

// march.metal:44-47.
static bool falls_in(const orc_ray& ray, int outcome) {
    return outcome == 0 || (outcome == 2 && ray.s.rdot < 0.0 && std::fabs(ray.angular_momentum) < B_CRITICAL * ray.energy);
}

static v3 sphere_exit_direction(const orc_march_params* p, const orc_ray& ray) {
    double b = std::fabs(ray.angular_momentum) / ray.energy;
    double sweep = sweep_lookup(p->sphere_table, 0.0, p->sphere_table_max_b, b);
    return asymptotic(ray, ray.s.phi + sweep);
}

static double kerr_step_length_kernel(double a, const orc_kerr_state& s, const orc_settings* g) {
    double near_h = 0.2 * (s.r - orc_kerr_outer_horizon(a));
    double pole = std::min(std::max(std::sin(s.theta) / KERR_POLE_SINE, 0.05), 1.0);
    return std::min(std::max(std::min(g->step_factor * s.r, near_h) * pole, KERR_STEP_FLOOR), g->step_max);
}

static double ksigma2(double a, double r, double theta) {
    double c = std::cos(theta);
    return r * r + a * a * c * c;
}

// march.metal:123-184.
static void march_kerr(const orc_march_params* p, const orc_field* field, orc_kerr_state& s, Gather& gather,
                       int& outcome, int64_t& steps, double& drift) {
    double a = p->spin;
    double carter0 = orc_kerr_carter(a, &s);
    double scale = std::max(std::fabs(carter0), s.energy * s.energy);
    double capture = orc_kerr_outer_horizon(a) + KERR_CAPTURE_MARGIN;
    double prev[3];
    orc_kerr_position(a, &s, prev);
    outcome = 2;
    steps = 0;
    drift = 0.0;
    const orc_settings* g = &p->geodesic;
    for (int64_t k = 0; k < g->step_cap; ++k) {
        orc_kerr_step(a, &s, kerr_step_length_kernel(a, s, g));
        steps = k + 1;
        double pos[3];
        orc_kerr_position(a, &s, pos);
        double path = orc::length(orc::make(pos) - orc::make(prev));
        gather.travelled += path;
        double carter_drift = std::fabs(orc_kerr_carter(a, &s) - carter0) / scale;
        double null_drift = std::fabs(orc_kerr_null_residual(a, &s)) / (scale * std::max(s.r * s.r, 1.0));
        drift = std::max(drift, std::max(carter_drift, null_drift));
        if (s.r <= capture) {
            outcome = 0;
            break;
        }
        double f = std::max(1.0 - 2.0 * s.r / ksigma2(a, s.r, s.theta), 0.0);
        double d[3];
        orc_kerr_direction(a, &s, d);
        gather.add(p, field, pos, path, f, d);
        prev[0] = pos[0];
        prev[1] = pos[1];
        prev[2] = pos[2];
        if (s.r >= g->escape_radius) {
            outcome = 1;
            break;
        }
    }
}

// march.metal:194-288 for one pixel.
void orc_march_pixel(const orc_march_params* p, const orc_field* field, int px, int py, orc_pixel* out) {
    double sx = (double)px + 0.5, sy = (double)py + 0.5;
    double nx = (2.0 * sx / (double)p->width - 1.0) * p->tan_half_fov[0];
    double ny = (1.0 - 2.0 * sy / (double)p->height) * p->tan_half_fov[1];
    v3 dir = orc::normalize(orc::make(p->camera_forward) + nx * orc::make(p->camera_right) + ny * orc::make(p->camera_up));
    double direction[3];
    orc::store(dir, direction);
    Gather gather;
    out->drift = 0.0;
    out->steps = 0;
    out->outcome = 1;
    if (p->spin > 0.0) {
        orc_kerr_state s;
        orc_kerr_launch(p->spin, p->camera_position, direction, &s);
        int outcome;
        int64_t steps;
        double drift;
        march_kerr(p, field, s, gather, outcome, steps, drift);
        out->drift = drift;
        out->steps = steps;
        out->outcome = outcome;
        out->falls_in = outcome == 0 || (outcome == 2 && s.pr < 0.0);
        double sky[3];
        orc_kerr_direction(p->spin, &s, sky);
        orc::store(orc::make(sky), out->sky);
        v3 color = gather.color;
        if (!out->falls_in && p->star_brightness > 0.0) {
            double star[3];
            orc_starfield(sky, p->star_seed, p->blackbody, star);
            color = color + (gather.transmittance * p->star_brightness) * orc::make(star);
        }
        orc::store(color, out->color);
        out->transmittance = gather.transmittance;
        double depth = gather.depth_weight > 0.0 ? gather.weighted_depth / gather.depth_weight : gather.travelled;
        out->depth = std::min(depth / (2.0 * R_ESCAPE), 1.0);
        return;
    }
    SchwMarch m = prepare(p, p->camera_position, direction);
    double depth = 2.0 * R_ESCAPE;
    v3 color = {0, 0, 0};
    out->falls_in = 0;
    if (m.sky_only) {
        orc::store(m.sky, out->sky);
        if (p->star_brightness > 0.0) {
            double star[3];
            orc_starfield(out->sky, p->star_seed, p->blackbody, star);
            color = p->star_brightness * orc::make(star);
        }
        out->transmittance = 1.0;
    } else {
        int outcome;
        int64_t steps;
        bool left;
        march_schw(p, field, m.ray, gather, outcome, steps, left);
        out->drift = schw_drift(m.ray);
        out->steps = steps;
        out->outcome = outcome;
        depth = gather.depth_weight > 0.0 ? gather.weighted_depth / gather.depth_weight : gather.travelled;
        color = gather.color;
        out->transmittance = gather.transmittance;
        out->falls_in = falls_in(m.ray, outcome);
        v3 sky;
        if (left) sky = sphere_exit_direction(p, m.ray);
        else {
            double d[3];
            orc_schw_direction(&m.ray, d);
            sky = orc::make(d);
        }
        orc::store(sky, out->sky);
        if (!out->falls_in && p->star_brightness > 0.0) {
            double star[3];
            orc_starfield(out->sky, p->star_seed, p->blackbody, star);
            color = color + (gather.transmittance * p->star_brightness) * orc::make(star);
        }
    }
    orc::store(color, out->color);
    out->depth = std::min(depth / (2.0 * R_ESCAPE), 1.0);
}

int64_t orc_march_pixels(const orc_march_params* p, const orc_field* field, const int32_t* xy, int64_t count,
                         int threads, orc_pixel* out) {
    std::atomic<int64_t> total{0};
    orc::parallel_dynamic(count, threads, 16, [&](int64_t b, int64_t e) {
        int64_t local = 0;
        for (int64_t i = b; i < e; ++i) {
            orc_march_pixel(p, field, xy[2 * i], xy[2 * i + 1], &out[i]);
            local += out[i].steps;
        }
        total += local;
    });
    return total.load();
}

// march.metal:290-341.
void orc_probe(const orc_march_params* p, const double origin[3], const double direction[3], double position_out[3],
               double direction_out[3], double* drift, int* outcome, int64_t* steps, double* energy,
               double* angular_momentum) {
    Gather none;
    if (p->spin > 0.0) {
        orc_kerr_state s;
        orc_kerr_launch(p->spin, origin, direction, &s);
        double d;
        march_kerr(p, nullptr, s, none, *outcome, *steps, d);
        *drift = d;
        orc_kerr_position(p->spin, &s, position_out);
        orc_kerr_direction(p->spin, &s, direction_out);
        *energy = s.energy;
        *angular_momentum = s.angular_momentum;
        return;
    }
    SchwMarch m = prepare(p, origin, direction);
    if (m.sky_only) {
        orc_schw_position(&m.ray, position_out);
        orc::store(m.sky, direction_out);
        *drift = 0.0;
        *outcome = 1;
        *steps = 0;
        *energy = m.ray.energy;
        *angular_momentum = m.ray.angular_momentum;
        return;
    }
    bool left;
    march_schw(p, nullptr, m.ray, none, *outcome, *steps, left);
    *drift = schw_drift(m.ray);
    orc_schw_position(&m.ray, position_out);
    if (*outcome == 1 && !falls_in(m.ray, *outcome)) {
        v3 sky;
        if (left) sky = sphere_exit_direction(p, m.ray);
        else {
            double d[3];
            orc_schw_direction(&m.ray, d);
            sky = orc::make(d);
        }
        orc::store(sky, direction_out);
    } else {
        orc_schw_direction(&m.ray, direction_out);
    }
    *energy = orc_schw_energy(&m.ray.s);
    *angular_momentum = orc_schw_angular_momentum(&m.ray.s);
}

// bake.metal:17-92: the march of marchRay without the gather, storing the
// in-slab points at least `spacing` of path apart.
void orc_bake_pixel(const orc_march_params* p, int px, int py, int64_t capacity, double spacing, double* samples,
                    orc_baked_ray* header) {
    double sx = (double)px + 0.5, sy = (double)py + 0.5;
    double nx = (2.0 * sx / (double)p->width - 1.0) * p->tan_half_fov[0];
    double ny = (1.0 - 2.0 * sy / (double)p->height) * p->tan_half_fov[1];
    v3 dir = orc::normalize(orc::make(p->camera_forward) + nx * orc::make(p->camera_right) + ny * orc::make(p->camera_up));
    double direction[3];
    orc::store(dir, direction);
    SchwMarch m = prepare(p, p->camera_position, direction);
    header->count = 0;
    header->falls_in = 0;
    header->overflowed = 0;
    if (m.sky_only) {
        orc::store(m.sky, header->direction);
        header->depth = 1.0;
        return;
    }
    orc_ray& ray = m.ray;
    const orc_settings* g = &p->geodesic;
    double prev[3];
    orc_schw_position(&ray, prev);
    double since = 0.0, travelled = 0.0, first_depth = -1.0;
    int outcome = 2;
    bool left = false;
    for (int64_t k = 0; k < g->step_cap; ++k) {
        double h = step_length(g, ray.s.r, prev[2]);
        orc_ray_state before = ray.s;
        orc_schw_step(&ray.s, h);
        if (p->sphere_radius > 0.0 && ray.s.r >= p->sphere_radius && ray.s.rdot > 0.0 && before.r < p->sphere_radius) {
            double frac = (p->sphere_radius - before.r) / (ray.s.r - before.r);
            ray.s = before;
            orc_schw_step(&ray.s, h * frac);
        }
        double pos[3];
        orc_schw_position(&ray, pos);
        double path = orc::length(orc::make(pos) - orc::make(prev));
        since += path;
        travelled += path;
        prev[0] = pos[0];
        prev[1] = pos[1];
        prev[2] = pos[2];
        if (ray.s.r <= g->capture_radius) {
            outcome = 0;
            break;
        }
        if (std::fabs(pos[2]) < DISK_SLAB_HALF_HEIGHT && std::fabs(pos[0]) < p->volume_half_extent &&
            std::fabs(pos[1]) < p->volume_half_extent && since >= spacing) {
            if (header->count < capacity) {
                double* s = samples + 4 * header->count;
                s[0] = pos[0];
                s[1] = pos[1];
                s[2] = pos[2];
                s[3] = since;
                header->count += 1;
                if (first_depth < 0.0) first_depth = travelled;
            } else {
                header->overflowed = 1;
            }
            since = 0.0;
        }
        if (p->sphere_radius > 0.0 && ray.s.r >= p->sphere_radius && ray.s.rdot > 0.0) {
            outcome = 1;
            left = true;
            break;
        }
        if (ray.s.r >= g->escape_radius) {
            outcome = 1;
            break;
        }
    }
    header->falls_in = falls_in(ray, outcome) ? 1 : 0;
    v3 fin;
    if (left) fin = sphere_exit_direction(p, ray);
    else {
        double d[3];
        orc_schw_direction(&ray, d);
        fin = orc::make(d);
    }
    orc::store(fin, header->direction);
    double depth = first_depth >= 0.0 ? first_depth : travelled;
    header->depth = std::min(depth / (2.0 * R_ESCAPE), 1.0);
}

// bake.metal:94-144: the stored samples against the current field, with the
// full Doppler factor cubed (no density blend) at f(|position|) along the
// chord from the previous sample, then the starfield.
void orc_walk_pixel(const orc_march_params* p, const orc_field* field, const double* samples, int64_t count,
                    int falls_in_flag, const double direction[3], double color_out[3]) {
    v3 color = {0, 0, 0};
    double transmittance = 1.0;
    v3 prev = orc::make(p->camera_position);
    for (int64_t k = 0; k < count; ++k) {
        if (transmittance <= TRANSMITTANCE_FLOOR) break;
        const double* s = samples + 4 * k;
        v3 pos = {s[0], s[1], s[2]};
        double path = s[3];
        if (field) {
            double e[4], v[3];
            orc_field_sample(field, s, e, v);
            if (e[3] > 0.0) {
                double g3 = 1.0;
                if (p->redshift) {
                    double d[3];
                    orc::store(orc::normalize(pos - prev), d);
                    double g = orc_redshift_factor(1.0 - R_SCHWARZSCHILD / orc::length(pos), d, v);
                    g3 = g * g * g;
                }
                double alpha = 1.0 - std::exp(-p->opacity_scale * e[3] * path);
                color = color + (transmittance * g3 * path) * v3{e[0], e[1], e[2]};
                transmittance *= 1.0 - alpha;
            }
        }
        prev = pos;
    }
    if (!falls_in_flag && p->star_brightness > 0.0) {
        double d[3], star[3];
        orc::store(orc::normalize(orc::make(direction)), d);
        orc_starfield(d, p->star_seed, p->blackbody, star);
        color = color + (transmittance * p->star_brightness) * orc::make(star);
    }
    orc::store(color, color_out);
}

}  // extern "C"
