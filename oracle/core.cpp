// FP64 oracle: hashing, potentials, particle dynamics, disk model, inner
// edge fit, blackbody colour. Test infrastructure only (see oracle.h).
#include "oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

extern "C" {

// Hash.swift:9-21; pcg4d of Jarzynski and Olano 2020, JCGT 9(3), 20.
void orc_pcg4d(const uint32_t in[4], uint32_t out[4]) {
    uint32_t v[4];
    for (int i = 0; i < 4; ++i) v[i] = in[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16u;
    v[0] += v[1] * v[3];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) out[i] = v[i];
}

// Hash.swift:25-29: lane selects a stream by xor into the low seed word.
void orc_hash_draw(uint64_t seed, uint32_t index, uint32_t event, uint32_t lane, uint32_t out[4]) {
    uint32_t low = (uint32_t)seed;
    uint32_t high = (uint32_t)(seed >> 32);
    uint32_t in[4] = {index, event, low ^ (lane * 0x9E3779B9u), high};
    orc_pcg4d(in, out);
}

// Hash.swift:32-34.
double orc_unit(uint32_t x) { return (double)(x >> 8) / 16777216.0; }

// Hash.swift:39-43, Box and Muller 1958.
void orc_gaussian_pair(double u1, double u2, double out[2]) {
    double radius = std::sqrt(-2.0 * std::log(1.0 - u1));
    double angle = 2.0 * M_PI * u2;
    out[0] = radius * std::cos(angle);
    out[1] = radius * std::sin(angle);
}

// Validate.swift:378-387, 64-bit FNV-1a (Fowler, Noll and Vo).
uint64_t orc_fnv1a(const void* bytes, uint64_t count) {
    const uint8_t* p = (const uint8_t*)bytes;
    uint64_t hash = 0xcbf29ce484222325ull;
    for (uint64_t i = 0; i < count; ++i) {
        hash ^= (uint64_t)p[i];
        hash *= 0x100000001b3ull;
    }
    return hash;
}

// Potential.swift:16-60.
static double potential_offset(double r) { return std::max(r - R_SCHWARZSCHILD, 1e-6); }

double orc_potential_energy(double r, int potential) {
    return potential == 1 ? -1.0 / potential_offset(r) : -1.0 / r;
}

double orc_potential_gradient(double r, int potential) {
    if (potential == 1) {
        double d = potential_offset(r);
        return 1.0 / (d * d);
    }
    return 1.0 / (r * r);
}

double orc_circular_speed(double r, int potential) { return std::sqrt(r * orc_potential_gradient(r, potential)); }
double orc_angular_frequency(double r, int potential) { return orc_circular_speed(r, potential) / r; }

double orc_circular_energy(double r, int potential) {
    double v = orc_circular_speed(r, potential);
    return 0.5 * v * v + orc_potential_energy(r, potential);
}

double orc_circular_angular_momentum(double r, int potential) { return r * orc_circular_speed(r, potential); }

// ParticleDynamics.swift:31-38: gravity kick, drag, drift.
void orc_particle_step(double s[6], int potential, double dt, double alpha) {
    double r = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    double k = -orc_potential_gradient(r, potential) / r;
    double omega = orc_angular_frequency(r, potential);
    for (int i = 0; i < 3; ++i) s[3 + i] += s[i] * k * dt;
    double drag = alpha * omega * dt;
    for (int i = 0; i < 3; ++i) s[3 + i] -= s[3 + i] * drag;
    for (int i = 0; i < 3; ++i) s[i] += s[3 + i] * dt;
}

// ParticleDynamics.swift:43-52: swept capture, escape, speed limit.
int orc_particle_removed(const double before[3], const double after[6]) {
    double travel[3] = {after[0] - before[0], after[1] - before[1], after[2] - before[2]};
    double length_squared = std::max(travel[0] * travel[0] + travel[1] * travel[1] + travel[2] * travel[2], 1e-12);
    double along = -(before[0] * travel[0] + before[1] * travel[1] + before[2] * travel[2]) / length_squared;
    along = std::min(std::max(along, 0.0), 1.0);
    double c[3] = {before[0] + along * travel[0], before[1] + along * travel[1], before[2] + along * travel[2]};
    double closest = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    double radius = std::sqrt(after[0] * after[0] + after[1] * after[1] + after[2] * after[2]);
    double speed = std::sqrt(after[3] * after[3] + after[4] * after[4] + after[5] * after[5]);
    return closest < R_CAPTURE || radius > R_ESCAPE || speed > PLUNGE_SPEED_LIMIT;
}

double orc_specific_energy(const double s[6], int potential) {
    double v2 = s[3] * s[3] + s[4] * s[4] + s[5] * s[5];
    double r = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    return 0.5 * v2 + orc_potential_energy(r, potential);
}

// DiskModel.swift:17-21.
double orc_disk_shape(double r, double floor) {
    double x = R_ISCO / r;
    double torque = std::max(1.0 - std::sqrt(x), floor);
    return std::pow(x, 0.75) * std::pow(torque, 0.25);
}

double orc_disk_temperature(double r, double peak) {
    double peak_radius = R_ISCO * 49.0 / 36.0;
    return peak * orc_disk_shape(r, ZERO_TORQUE_FLOOR) / orc_disk_shape(peak_radius, ZERO_TORQUE_FLOOR);
}

// DiskModel.swift:31-35.
static double mass_flux_fraction(double r) {
    if (r <= R_FEED_INNER) return 1.0;
    if (r >= R_DISK_OUTER) return 0.0;
    return (R_DISK_OUTER * R_DISK_OUTER - r * r) / (R_DISK_OUTER * R_DISK_OUTER - R_FEED_INNER * R_FEED_INNER);
}

double orc_steady_surface_density(double r) {
    if (r <= R_ISCO || r >= R_DISK_OUTER) return 0.0;
    return mass_flux_fraction(r) * (r - R_ISCO) / std::pow(r, 1.5);
}

double orc_radial_number_density(double r) { return 2.0 * M_PI * r * orc_steady_surface_density(r); }

// DiskModel.swift:54-78: trapezoid cumulative, linear inversion.
void orc_radial_inverse_cdf(int count, int resolution, double* table) {
    double r0 = R_ISCO, r1 = R_DISK_OUTER;
    double dr = (r1 - r0) / (double)resolution;
    std::vector<double> cumulative(resolution + 1, 0.0);
    double previous = orc_radial_number_density(r0);
    for (int i = 1; i <= resolution; ++i) {
        double r = r0 + (double)i * dr;
        double density = orc_radial_number_density(r);
        cumulative[i] = cumulative[i - 1] + 0.5 * (previous + density) * dr;
        previous = density;
    }
    double total = cumulative[resolution];
    for (int k = 0; k < count; ++k) table[k] = r0;
    int i = 0;
    for (int k = 0; k < count; ++k) {
        double target = total * (double)k / (double)(count - 1);
        while (i < resolution && cumulative[i + 1] < target) ++i;
        if (i >= resolution) {
            table[k] = r1;
            continue;
        }
        double span = cumulative[i + 1] - cumulative[i];
        double t = span > 0 ? (target - cumulative[i]) / span : 0.0;
        table[k] = r0 + ((double)i + t) * dr;
    }
    table[count - 1] = r1;
}

// DiskModel.swift:83-107 (generator 0); generator 1 is the annulus scene
// generator of the benchmark configurations: the respawn radius law at every
// event and a fixed Gaussian height sigma_z.
void orc_spawn(uint32_t index, uint32_t event, uint64_t seed, const double* cdf, int cdf_count, int potential,
               int generator, double sigma_z, double out[6]) {
    uint32_t h0[4], h1[4];
    orc_hash_draw(seed, index, event, 0, h0);
    orc_hash_draw(seed, index, event, 1, h1);
    double u0 = orc_unit(h0[0]), u1 = orc_unit(h0[1]), u2 = orc_unit(h0[2]), u3 = orc_unit(h0[3]);
    double u4 = orc_unit(h1[0]), u5 = orc_unit(h1[1]);
    double r;
    if (event == 0 && generator == 0) {
        double position = u0 * (double)(cdf_count - 1);
        int k = std::min((int)position, cdf_count - 2);
        r = cdf[k] + (position - (double)k) * (cdf[k + 1] - cdf[k]);
    } else {
        double inner = R_FEED_INNER * R_FEED_INNER;
        double outer = R_DISK_OUTER * R_DISK_OUTER;
        r = std::sqrt(inner + (outer - inner) * u0);
    }
    double phi = 2.0 * M_PI * u1;
    double g01[2], g23[2];
    orc_gaussian_pair(u2, u3, g01);
    orc_gaussian_pair(u4, u5, g23);
    double vc = orc_circular_speed(r, potential);
    double height = (generator == 1 && event == 0) ? sigma_z * g01[0] : DISK_ASPECT * r * g01[0];
    out[0] = r * std::cos(phi);
    out[1] = r * std::sin(phi);
    out[2] = height;
    out[3] = -vc * std::sin(phi) + FEED_VELOCITY_DISPERSION * vc * g01[1];
    out[4] = vc * std::cos(phi) + FEED_VELOCITY_DISPERSION * vc * g23[0];
    out[5] = FEED_VELOCITY_DISPERSION * vc * g23[1];
}

// DiskEdge.swift:28-43.
double orc_inner_edge(const double* radii, int64_t n, double bin_width, double fit_inner, double fit_outer) {
    int count = (int)std::llround((fit_outer - fit_inner) / bin_width);
    std::vector<double> counts(count, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        double r = radii[i];
        if (r >= fit_inner && r < fit_outer) counts[std::min((int)((r - fit_inner) / bin_width), count - 1)] += 1.0;
    }
    double sx = 0, sy = 0, sxx = 0, sxy = 0;
    for (int k = 0; k < count; ++k) {
        double center = fit_inner + ((double)k + 0.5) * bin_width;
        double sigma = counts[k] / (2.0 * M_PI * center * bin_width);
        double y = sigma * std::pow(center, 1.5);
        sx += center;
        sy += y;
        sxx += center * center;
        sxy += center * y;
    }
    double m = (double)count;
    double slope = (m * sxy - sx * sy) / (m * sxx - sx * sx);
    double intercept = (sy - slope * sx) / m;
    return -intercept / slope;
}

// GoldenTests.swift:39-64: seed, step, respawn on removal, cylindrical radii.
void orc_disk_run(uint32_t count, uint64_t seed, int substeps, int threads, double* radii, uint64_t* respawns_out) {
    std::vector<double> cdf(INVERSE_CDF_COUNT);
    orc_radial_inverse_cdf((int)INVERSE_CDF_COUNT, (int)INVERSE_CDF_RESOLUTION, cdf.data());
    std::vector<uint64_t> respawned(std::max(threads, 1), 0);
    orc::parallel_for(count, threads, [&](int64_t begin, int64_t end, int worker) {
        for (int64_t i = begin; i < end; ++i) {
            double s[6];
            uint32_t event = 0;
            orc_spawn((uint32_t)i, 0, seed, cdf.data(), (int)cdf.size(), 1, 0, 0.0, s);
            for (int k = 0; k < substeps; ++k) {
                double before[3] = {s[0], s[1], s[2]};
                orc_particle_step(s, 1, SIM_DT, DRAG_ALPHA);
                if (orc_particle_removed(before, s)) {
                    ++event;
                    respawned[worker] += 1;
                    orc_spawn((uint32_t)i, event, seed, cdf.data(), (int)cdf.size(), 1, 0, 0.0, s);
                }
            }
            radii[i] = std::sqrt(s[0] * s[0] + s[1] * s[1]);
        }
    });
    uint64_t total = 0;
    for (uint64_t r : respawned) total += r;
    if (respawns_out) *respawns_out = total;
}

// Blackbody.swift:13-82: Planck spectrum, Wyman, Sloan and Shirley 2013
// multi-lobe CIE fit, IEC 61966-2-1 linear sRGB.
static double planck(double nm, double temperature) {
    const double c2 = 1.438776877e-2;
    double lambda = nm * 1e-9;
    return 1.0 / (std::pow(lambda, 5.0) * (std::exp(c2 / (lambda * temperature)) - 1.0));
}

static double lobe(double x, double mu, double s1, double s2) {
    double sigma = x < mu ? s1 : s2;
    double t = (x - mu) / sigma;
    return std::exp(-0.5 * t * t);
}

void orc_blackbody_xyz(double temperature, double out[3]) {
    double sum[3] = {0, 0, 0};
    for (double w = 380.0; w <= 780.0; w += 1.0) {
        double xb = 1.056 * lobe(w, 599.8, 37.9, 31.0) + 0.362 * lobe(w, 442.0, 16.0, 26.7) - 0.065 * lobe(w, 501.1, 20.4, 26.2);
        double yb = 0.821 * lobe(w, 568.8, 46.9, 40.5) + 0.286 * lobe(w, 530.9, 16.3, 31.1);
        double zb = 1.217 * lobe(w, 437.0, 11.8, 36.0) + 0.681 * lobe(w, 459.0, 26.0, 13.4);
        double p = planck(w, temperature);
        sum[0] += xb * p;
        sum[1] += yb * p;
        sum[2] += zb * p;
    }
    out[0] = sum[0];
    out[1] = sum[1];
    out[2] = sum[2];
}

void orc_blackbody_chromaticity(double temperature, double out[2]) {
    double c[3];
    orc_blackbody_xyz(temperature, c);
    double total = c[0] + c[1] + c[2];
    out[0] = c[0] / total;
    out[1] = c[1] / total;
}

void orc_blackbody_color(double temperature, double out[3]) {
    double c[3];
    orc_blackbody_xyz(temperature, c);
    double rgb[3] = {3.2406 * c[0] - 1.5372 * c[1] - 0.4986 * c[2],
                     -0.9689 * c[0] + 1.8758 * c[1] + 0.0415 * c[2],
                     0.0557 * c[0] - 0.2040 * c[1] + 1.0570 * c[2]};
    for (int i = 0; i < 3; ++i) rgb[i] = std::max(rgb[i], 0.0);
    double peak = std::max(rgb[0], std::max(rgb[1], rgb[2]));
    for (int i = 0; i < 3; ++i) out[i] = peak > 0.0 ? rgb[i] / peak : 1.0;
}

void orc_blackbody_table(int count, double tmin, double tmax, double* out) {
    for (int k = 0; k < count; ++k) {
        double t = (double)k / (double)(count - 1);
        double temperature = std::exp(std::log(tmin) + t * (std::log(tmax) - std::log(tmin)));
        orc_blackbody_color(temperature, out + 3 * k);
    }
}

double orc_blackbody_coordinate(double temperature) {
    double t = std::min(std::max(temperature, BLACKBODY_T_MIN), BLACKBODY_T_MAX);
    return (std::log(t) - std::log(BLACKBODY_T_MIN)) / (std::log(BLACKBODY_T_MAX) - std::log(BLACKBODY_T_MIN));
}

// present.metal:19-35, Khronos PBR Neutral tone mapper.
void orc_pbr_neutral(const double in[3], double out[3]) {
    const double start = 0.8 - 0.04, desaturation = 0.15;
    double c[3] = {in[0], in[1], in[2]};
    double x = std::min(c[0], std::min(c[1], c[2]));
    double offset = x < 0.08 ? x - 6.25 * x * x : 0.04;
    for (int i = 0; i < 3; ++i) c[i] -= offset;
    double peak = std::max(c[0], std::max(c[1], c[2]));
    if (peak < start) {
        for (int i = 0; i < 3; ++i) out[i] = c[i];
        return;
    }
    double d = 1.0 - start;
    double new_peak = 1.0 - d * d / (peak + d - start);
    for (int i = 0; i < 3; ++i) c[i] *= new_peak / peak;
    double g = 1.0 - 1.0 / (desaturation * (peak - new_peak) + 1.0);
    for (int i = 0; i < 3; ++i) out[i] = c[i] + (new_peak - c[i]) * g;
}

// IEC 61966-2-1 sRGB transfer function (what a bgra8Unorm_srgb target applies).
double orc_srgb_encode(double x) {
    x = std::min(std::max(x, 0.0), 1.0);
    return x <= 0.0031308 ? 12.92 * x : 1.055 * std::pow(x, 1.0 / 2.4) - 0.055;
}

}  // extern "C"
