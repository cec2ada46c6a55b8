// Host-built FP64 tables the kernels read, made once per context the way the
// reference builds them at startup (Sim/ParticleSystem.swift:17-48): the
// blackbody colour table (Oracle/Blackbody.swift:13-82) and the inverse
// CDF of the steady surface density profile (Oracle/DiskModel.swift:31-78).
#include "host_tables.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/gargantua_constants.h"

namespace garg {
namespace {

// Planck radiance up to a constant, wavelength in nm; c2 = hc/k (CODATA 2018).
double planck(double nm, double kelvin) {
    const double c2 = 1.438776877e-2;
    double lambda = nm * 1e-9;
    return 1.0 / (std::pow(lambda, 5.0) * std::expm1(c2 / (lambda * kelvin)));
}

// Piecewise Gaussian lobe of the Wyman, Sloan and Shirley 2013 CIE fit.
double lobe(double x, double mu, double below, double above) {
    double t = (x - mu) / (x < mu ? below : above);
    return std::exp(-0.5 * t * t);
}

void xyz_of(double kelvin, double out[3]) {
    double x = 0, y = 0, z = 0;
    for (int nm = 380; nm <= 780; ++nm) {
        double w = (double)nm;
        double p = planck(w, kelvin);
        x += p * (1.056 * lobe(w, 599.8, 37.9, 31.0) + 0.362 * lobe(w, 442.0, 16.0, 26.7) -
                  0.065 * lobe(w, 501.1, 20.4, 26.2));
        y += p * (0.821 * lobe(w, 568.8, 46.9, 40.5) + 0.286 * lobe(w, 530.9, 16.3, 31.1));
        z += p * (1.217 * lobe(w, 437.0, 11.8, 36.0) + 0.681 * lobe(w, 459.0, 26.0, 13.4));
    }
    out[0] = x;
    out[1] = y;
    out[2] = z;
}

double surface_density(double r) {
    if (r <= R_ISCO || r >= R_DISK_OUTER) return 0.0;
    double flux = r <= R_FEED_INNER ? 1.0
                                    : (R_DISK_OUTER * R_DISK_OUTER - r * r) /
                                          (R_DISK_OUTER * R_DISK_OUTER - R_FEED_INNER * R_FEED_INNER);
    return flux * (r - R_ISCO) / std::pow(r, 1.5);
}

}  // namespace

void build_blackbody_table(std::vector<float>& rgba) {
    const int n = (int)BLACKBODY_TABLE_SIZE;
    rgba.assign(4 * n, 0.0f);
    const double lo = std::log(BLACKBODY_T_MIN), hi = std::log(BLACKBODY_T_MAX);
    for (int k = 0; k < n; ++k) {
        double kelvin = std::exp(lo + (double)k / (double)(n - 1) * (hi - lo));
        double c[3];
        xyz_of(kelvin, c);
        // XYZ to linear sRGB, D65 (IEC 61966-2-1), clipped and peak-normalized.
        double rgb[3] = {3.2406 * c[0] - 1.5372 * c[1] - 0.4986 * c[2],
                         -0.9689 * c[0] + 1.8758 * c[1] + 0.0415 * c[2],
                         0.0557 * c[0] - 0.2040 * c[1] + 1.0570 * c[2]};
        double peak = 0.0;
        for (double& v : rgb) {
            v = std::max(v, 0.0);
            peak = std::max(peak, v);
        }
        for (int i = 0; i < 3; ++i) rgba[4 * k + i] = (float)(peak > 0.0 ? rgb[i] / peak : 1.0);
        rgba[4 * k + 3] = 1.0f;
    }
}

void build_inverse_cdf(std::vector<float>& table) {
    const int count = (int)INVERSE_CDF_COUNT, resolution = (int)INVERSE_CDF_RESOLUTION;
    const double r0 = R_ISCO, r1 = R_DISK_OUTER, dr = (r1 - r0) / resolution;
    // Trapezoid cumulative of 2 pi r Sigma(r), then linear inversion.
    std::vector<double> cumulative(resolution + 1, 0.0);
    double previous = 2.0 * M_PI * r0 * surface_density(r0);
    for (int i = 1; i <= resolution; ++i) {
        double r = r0 + i * dr;
        double density = 2.0 * M_PI * r * surface_density(r);
        cumulative[i] = cumulative[i - 1] + 0.5 * (previous + density) * dr;
        previous = density;
    }
    table.assign(count, (float)r0);
    int i = 0;
    for (int k = 0; k < count; ++k) {
        double target = cumulative[resolution] * (double)k / (double)(count - 1);
        while (i < resolution && cumulative[i + 1] < target) ++i;
        if (i >= resolution) {
            table[k] = (float)r1;
            continue;
        }
        double span = cumulative[i + 1] - cumulative[i];
        double t = span > 0.0 ? (target - cumulative[i]) / span : 0.0;
        table[k] = (float)(r0 + (i + t) * dr);
    }
    table[count - 1] = (float)r1;
}

double temperature_scale() {
    // T_peak over the unnormalized profile at its peak, r = 49/36 r_isco
    // (Sim/ParticleSystem.swift:59-70).
    double x = 36.0 / 49.0;
    double shape = std::pow(x, 0.75) * std::pow(std::max(1.0 - std::sqrt(x), ZERO_TORQUE_FLOOR), 0.25);
    return T_PEAK_DEFAULT / shape;
}

}  // namespace garg
