// FP64 oracle: Schwarzschild null geodesics in ingoing Eddington-Finkelstein
// coordinates, far-field sweep tables, Kerr null geodesics in Boyer-Lindquist
// coordinates. Test infrastructure only (see oracle.h).
#include "oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>

using orc::v3;

namespace orc {
void parallel_dynamic(int64_t n, int threads, int64_t block, const std::function<void(int64_t, int64_t)>& body) {
    std::atomic<int64_t> next{0};
    auto worker = [&]() {
        for (;;) {
            int64_t begin = next.fetch_add(block);
            if (begin >= n) break;
            body(begin, std::min(n, begin + block));
        }
    };
    if (threads <= 1) {
        worker();
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
}
}  // namespace orc

extern "C" {

// Schwarzschild.swift:149-172 (Settings.interactive / .still).
void orc_settings_interactive(orc_settings* s) {
    s->step_factor = STEP_RADIUS_FACTOR;
    s->step_min = STEP_MIN;
    s->step_max = STEP_MAX;
    s->capture_radius = R_CAPTURE;
    s->escape_radius = R_ESCAPE;
    s->step_cap = STEP_CAP_INTERACTIVE;
    s->empty_step_factor = STEP_FACTOR_EMPTY_INTERACTIVE;
    s->empty_step_max = STEP_MAX_EMPTY_INTERACTIVE;
}

void orc_settings_still(orc_settings* s) {
    orc_settings_interactive(s);
    s->step_cap = STEP_CAP_STILL;
    s->empty_step_factor = STEP_FACTOR_EMPTY_STILL;
    s->empty_step_max = STEP_MAX_EMPTY_STILL;
}

double orc_schw_f(double r) { return 1.0 - R_SCHWARZSCHILD / r; }

double orc_schw_energy(const orc_ray_state* s) { return orc_schw_f(s->r) * s->vdot - s->rdot; }
double orc_schw_angular_momentum(const orc_ray_state* s) { return s->r * s->r * s->phidot; }

// Schwarzschild.swift:115-117.
double orc_schw_null_residual(const orc_ray_state* s) {
    return -orc_schw_f(s->r) * s->vdot * s->vdot + 2.0 * s->vdot * s->rdot + s->r * s->r * s->phidot * s->phidot;
}

// Schwarzschild.swift:76-97: static observer launch, local energy one.
void orc_schw_launch(const double origin_in[3], const double direction_in[3], orc_ray* out) {
    v3 origin = orc::make(origin_in), d = orc::make(direction_in);
    double r0 = orc::length(origin);
    v3 e1 = origin / r0;
    v3 unit = orc::normalize(d);
    double dr = orc::dot(unit, e1);
    v3 perpendicular = unit - dr * e1;
    double dphi = orc::length(perpendicular);
    if (dphi < 1e-12) {
        v3 helper = std::fabs(e1.x) < 0.9 ? v3{1.0, 0.0, 0.0} : v3{0.0, 1.0, 0.0};
        perpendicular = helper - orc::dot(helper, e1) * e1;
        perpendicular = perpendicular / orc::length(perpendicular);
        dphi = 0.0;
    } else {
        perpendicular = perpendicular / dphi;
    }
    double f0 = orc_schw_f(r0);
    out->s = {0.0, r0, 0.0, (1.0 + dr) / std::sqrt(f0), std::sqrt(f0) * dr, dphi / r0};
    orc::store(e1, out->e1);
    orc::store(perpendicular, out->e2);
    out->energy = orc_schw_energy(&out->s);
    out->angular_momentum = orc_schw_angular_momentum(&out->s);
}

// Schwarzschild.swift:100-104.
void orc_schw_launch_b(double radius, double b, orc_ray* out) {
    double dphi = b * std::sqrt(orc_schw_f(radius)) / radius;
    double dr = -std::sqrt(std::max(0.0, 1.0 - dphi * dphi));
    double origin[3] = {radius, 0.0, 0.0}, direction[3] = {dr, dphi, 0.0};
    orc_schw_launch(origin, direction, out);
}

// Schwarzschild.swift:121-127.
static orc_ray_state derivative(const orc_ray_state& s) {
    double r = s.r;
    double vdd = -s.vdot * s.vdot / (r * r) + r * s.phidot * s.phidot;
    double rdd = 2.0 * s.rdot * s.vdot / (r * r) + orc_schw_f(r) * (r * s.phidot * s.phidot - s.vdot * s.vdot / (r * r));
    double pdd = -2.0 * s.rdot * s.phidot / r;
    return {s.vdot, s.rdot, s.phidot, vdd, rdd, pdd};
}

static orc_ray_state axpy(const orc_ray_state& s, const orc_ray_state& k, double h) {
    return {s.v + k.v * h, s.r + k.r * h, s.phi + k.phi * h, s.vdot + k.vdot * h, s.rdot + k.rdot * h, s.phidot + k.phidot * h};
}

// Schwarzschild.swift:130-136, classical RK4.
void orc_schw_step(orc_ray_state* s, double h) {
    orc_ray_state k1 = derivative(*s);
    orc_ray_state k2 = derivative(axpy(*s, k1, 0.5 * h));
    orc_ray_state k3 = derivative(axpy(*s, k2, 0.5 * h));
    orc_ray_state k4 = derivative(axpy(*s, k3, h));
    double w = h / 6.0;
    s->v += (k1.v + k2.v * 2.0 + k3.v * 2.0 + k4.v) * w;
    s->r += (k1.r + k2.r * 2.0 + k3.r * 2.0 + k4.r) * w;
    s->phi += (k1.phi + k2.phi * 2.0 + k3.phi * 2.0 + k4.phi) * w;
    s->vdot += (k1.vdot + k2.vdot * 2.0 + k3.vdot * 2.0 + k4.vdot) * w;
    s->rdot += (k1.rdot + k2.rdot * 2.0 + k3.rdot * 2.0 + k4.rdot) * w;
    s->phidot += (k1.phidot + k2.phidot * 2.0 + k3.phidot * 2.0 + k4.phidot) * w;
}

void orc_schw_position(const orc_ray* ray, double out[3]) {
    v3 e1 = orc::make(ray->e1), e2 = orc::make(ray->e2);
    orc::store(ray->s.r * (std::cos(ray->s.phi) * e1 + std::sin(ray->s.phi) * e2), out);
}

// Schwarzschild.swift:61-67.
void orc_schw_direction(const orc_ray* ray, double out[3]) {
    v3 e1 = orc::make(ray->e1), e2 = orc::make(ray->e2);
    double c = std::cos(ray->s.phi), s = std::sin(ray->s.phi);
    v3 radial = c * e1 + s * e2;
    v3 tangential = -s * e1 + c * e2;
    v3 local = ray->s.rdot / std::sqrt(std::max(orc_schw_f(ray->s.r), 1e-6)) * radial + ray->s.r * ray->s.phidot * tangential;
    double len = orc::length(local);
    orc::store(len > 0 ? local / len : radial, out);
}

// Schwarzschild.swift:174-184: the empty-leg coarse step applies when the ray
// is farther from the slab than that step.
static double step_length(const orc_settings* g, double r, double z) {
    double coarse = std::min(std::max(g->empty_step_factor * r, g->step_min), g->empty_step_max);
    if (std::fabs(z) > DISK_SLAB_HALF_HEIGHT + coarse) return coarse;
    return std::min(std::max(g->step_factor * r, g->step_min), g->step_max);
}

// Schwarzschild.swift:200-220.
void orc_schw_integrate(const orc_ray* launch, const orc_settings* settings, orc_result* out) {
    orc_ray ray = *launch;
    int64_t steps = 0;
    double drift = 0.0;
    double l_scale = std::max(std::fabs(launch->angular_momentum), 1e-3);
    int outcome = 2;
    while (steps < settings->step_cap) {
        double p[3];
        orc_schw_position(&ray, p);
        orc_schw_step(&ray.s, step_length(settings, ray.s.r, p[2]));
        ++steps;
        double e_drift = std::fabs(orc_schw_energy(&ray.s) - launch->energy) / std::fabs(launch->energy);
        double l_drift = std::fabs(orc_schw_angular_momentum(&ray.s) - launch->angular_momentum) / l_scale;
        drift = std::max(drift, std::max(e_drift, l_drift));
        if (ray.s.r <= settings->capture_radius) {
            outcome = 0;
            break;
        }
        if (ray.s.r >= settings->escape_radius) {
            outcome = 1;
            break;
        }
    }
    out->ray = ray;
    out->outcome = outcome;
    out->steps = steps;
    out->drift = drift;
}

// Schwarzschild.swift:244-252.
double orc_schw_deflection(const orc_ray* launch, const orc_result* result) {
    double a3[3], b3[3];
    orc_schw_direction(launch, a3);
    orc_schw_direction(&result->ray, b3);
    double ax = a3[0] * launch->e1[0] + a3[1] * launch->e1[1] + a3[2] * launch->e1[2];
    double ay = a3[0] * launch->e2[0] + a3[1] * launch->e2[1] + a3[2] * launch->e2[2];
    double bx = b3[0] * result->ray.e1[0] + b3[1] * result->ray.e1[1] + b3[2] * result->ray.e1[2];
    double by = b3[0] * result->ray.e2[0] + b3[1] * result->ray.e2[1] + b3[2] * result->ray.e2[2];
    return std::atan2(ax * by - ay * bx, ax * bx + ay * by);
}

// Schwarzschild.swift:256-269.
double orc_schw_critical_b(double launch_radius, const orc_settings* settings, int iterations) {
    double low = 3.0, high = 8.0;
    for (int i = 0; i < iterations; ++i) {
        double mid = 0.5 * (low + high);
        orc_ray ray;
        orc_schw_launch_b(launch_radius, mid, &ray);
        orc_result result;
        orc_schw_integrate(&ray, settings, &result);
        if (result.outcome == 0) low = mid;
        else if (result.outcome == 1) high = mid;
        else return mid;
    }
    return 0.5 * (low + high);
}

// Schwarzschild.swift:230-240 with f given directly (Geodesic.h:225-236).
double orc_redshift_factor(double f, const double direction[3], const double gas[3]) {
    v3 g = orc::make(gas);
    double magnitude = orc::length(g);
    double speed = std::min(magnitude, 8.0);
    double beta = speed > 0 ? GAS_SPEED_CEILING * std::tanh(speed / GAS_SPEED_CEILING) : 0.0;
    v3 velocity = magnitude > 0 ? g * (beta / magnitude) : v3{0, 0, 0};
    double gamma = 1.0 / std::sqrt(1.0 - beta * beta);
    return std::sqrt(std::max(f, 0.0)) / (gamma * (1.0 + orc::dot(orc::make(direction), velocity)));
}

// ---- Deflection.swift ----------------------------------------------------

double orc_max_impact_parameter(double r) { return r / std::sqrt(1.0 - R_SCHWARZSCHILD / r); }

// Deflection.swift:23-39: midpoint rule after u = uMax sin^2(theta).
double orc_sweep(double r, double b, int intervals) {
    if (!(b > 0.0)) return 0.0;
    double bound = std::min(b, orc_max_impact_parameter(r));
    double u_max = 1.0 / r;
    double h = (M_PI / 2.0) / (double)intervals;
    double sum = 0.0;
    for (int k = 0; k < intervals; ++k) {
        double theta = ((double)k + 0.5) * h;
        double s = std::sin(theta), c = std::cos(theta);
        double u = u_max * s * s;
        double inner = 1.0 - bound * bound * u * u * (1.0 - 2.0 * u);
        if (inner > 0.0) sum += bound * 2.0 * u_max * s * c / std::sqrt(inner);
    }
    return sum * h;
}

// Deflection.swift:45-54.
double orc_table_b(double low, double high, double t) {
    double back = 1.0 - t;
    return low + (high - low) * (1.0 - back * back);
}

double orc_table_coordinate(double low, double high, double b) {
    double x = std::min(std::max((b - low) / (high - low), 0.0), 1.0);
    return 1.0 - std::sqrt(1.0 - x);
}

// Deflection.swift:59-69.
double orc_periapsis(double b) {
    double u = 1.0 / b;
    for (int i = 0; i < 60; ++i) {
        double value = 1.0 - b * b * u * u * (1.0 - 2.0 * u);
        double slope = -b * b * (2.0 * u - 6.0 * u * u);
        double next = u - value / slope;
        if (std::fabs(next - u) < 1e-15) {
            u = next;
            break;
        }
        u = next;
    }
    return 1.0 / u;
}

// Deflection.swift:77-86.
double orc_turning_sweep(double r, double b, int intervals) {
    double peri = orc_periapsis(b);
    return 2.0 * (orc_sweep(peri, b, intervals) - orc_sweep(r, b, intervals));
}

double orc_escape_sweep(double r, double b, int intervals) {
    return orc_turning_sweep(r, b, intervals) + orc_sweep(r, b, intervals);
}

// Deflection.swift:89-100.
void orc_sweep_table(double r, int count, int intervals, double* out) {
    double b_max = orc_max_impact_parameter(r);
    for (int i = 0; i < count; ++i) out[i] = orc_sweep(r, orc_table_b(0.0, b_max, (double)i / (double)(count - 1)), intervals);
}

void orc_sky_table(double camera_r, double sphere_r, int count, int intervals, double* out) {
    double low = orc_max_impact_parameter(sphere_r), high = orc_max_impact_parameter(camera_r);
    for (int i = 0; i < count; ++i)
        out[i] = orc_escape_sweep(camera_r, orc_table_b(low, high, (double)i / (double)(count - 1)), intervals);
}

void orc_far_field_table(double r_obs, int count, double b_lo, double b_hi, int intervals, double* b_out,
                         double* alpha_out, double* sweep_out) {
    for (int k = 0; k < count; ++k) {
        double b = b_lo * std::pow(b_hi / b_lo, (double)k / (double)(count - 1));
        b_out[k] = b;
        alpha_out[k] = 2.0 * orc_sweep(orc_periapsis(b), b, intervals) - M_PI;
        sweep_out[k] = orc_sweep(r_obs, b, intervals);
    }
}

// Bozza 2002, Phys. Rev. D 66, 103001, eq. (35) for Schwarzschild with
// a-bar = 1, b-bar = -pi + ln[216 (7 - 4 sqrt 3)], u_m = b_c.
double orc_strong_deflection(double b) {
    return -std::log(b / B_CRITICAL - 1.0) + std::log(216.0 * (7.0 - 4.0 * std::sqrt(3.0))) - M_PI;
}

// ---- KerrGeodesic.swift ---------------------------------------------------

// KerrGeodesic.swift:176-197.
void orc_kerr_settings_default(orc_kerr_settings* s, int64_t step_cap) {
    s->step_factor = STEP_RADIUS_FACTOR;
    s->step_min = 0.005;
    s->step_max = STEP_MAX;
    s->capture_margin = KERR_CAPTURE_MARGIN;
    s->escape_radius = R_ESCAPE;
    s->step_cap = step_cap;
}

double orc_kerr_outer_horizon(double a) { return 1.0 + std::sqrt(1.0 - a * a); }

static double kdelta(double a, double r) { return r * r - 2.0 * r + a * a; }
static double ksigma(double a, double r, double theta) {
    double c = std::cos(theta);
    return r * r + a * a * c * c;
}

// KerrGeodesic.swift:52-59.
static void potentials(double a, const orc_kerr_state& s, double& p, double& g) {
    double e = s.energy, l = s.angular_momentum;
    p = (s.r * s.r + a * a) * e - a * l;
    double sin2 = std::max(std::sin(s.theta) * std::sin(s.theta), 1e-12);
    g = l * l / sin2 - 2.0 * a * e * l + a * a * e * e * sin2;
}

double orc_kerr_null_residual(double a, const orc_kerr_state* s) {
    double delta = kdelta(a, s->r), p, g;
    potentials(a, *s, p, g);
    return delta * s->pr * s->pr + s->ptheta * s->ptheta - p * p / delta + g;
}

double orc_kerr_carter(double a, const orc_kerr_state* s) {
    double c = std::cos(s->theta);
    double sin2 = std::max(std::sin(s->theta) * std::sin(s->theta), 1e-12);
    double term = s->angular_momentum * s->angular_momentum / sin2 - a * a * s->energy * s->energy;
    return s->ptheta * s->ptheta + c * c * term;
}

// KerrGeodesic.swift:77-105.
static orc_kerr_state kderivative(double a, const orc_kerr_state& s) {
    double e = s.energy, l = s.angular_momentum;
    double r = s.r, theta = s.theta;
    double delta = kdelta(a, r), sigma = ksigma(a, r, theta);
    double delta_prime = 2.0 * r - 2.0;
    double st = std::sin(theta), ct = std::cos(theta);
    double sin2 = std::max(st * st, 1e-12);
    double p, g;
    potentials(a, s, p, g);
    double f = delta * s.pr * s.pr + s.ptheta * s.ptheta - p * p / delta + g;
    double hamiltonian = f / (2.0 * sigma);
    double p_prime = 2.0 * r * e;
    double fr = delta_prime * s.pr * s.pr - 2.0 * p * p_prime / delta + p * p * delta_prime / (delta * delta);
    double ftheta = -2.0 * l * l * ct / (sin2 * std::max(st, 1e-9)) + 2.0 * a * a * e * e * st * ct;
    double sigma_r = 2.0 * r, sigma_theta = -2.0 * a * a * st * ct;
    orc_kerr_state out = s;
    out.r = delta * s.pr / sigma;
    out.theta = s.ptheta / sigma;
    out.phi = (a * p / delta + l / sin2 - a * e) / sigma;
    out.pr = -(fr / (2.0 * sigma) - hamiltonian * sigma_r / sigma);
    out.ptheta = -(ftheta / (2.0 * sigma) - hamiltonian * sigma_theta / sigma);
    return out;
}

static orc_kerr_state kaxpy(const orc_kerr_state& s, const orc_kerr_state& k, double h) {
    orc_kerr_state o = s;
    o.r += k.r * h;
    o.theta += k.theta * h;
    o.phi += k.phi * h;
    o.pr += k.pr * h;
    o.ptheta += k.ptheta * h;
    return o;
}

// KerrGeodesic.swift:107-113.
void orc_kerr_step(double a, orc_kerr_state* s, double h) {
    orc_kerr_state k1 = kderivative(a, *s);
    orc_kerr_state k2 = kderivative(a, kaxpy(*s, k1, 0.5 * h));
    orc_kerr_state k3 = kderivative(a, kaxpy(*s, k2, 0.5 * h));
    orc_kerr_state k4 = kderivative(a, kaxpy(*s, k3, h));
    double w = h / 6.0;
    s->r += (k1.r + k2.r * 2.0 + k3.r * 2.0 + k4.r) * w;
    s->theta += (k1.theta + k2.theta * 2.0 + k3.theta * 2.0 + k4.theta) * w;
    s->phi += (k1.phi + k2.phi * 2.0 + k3.phi * 2.0 + k4.phi) * w;
    s->pr += (k1.pr + k2.pr * 2.0 + k3.pr * 2.0 + k4.pr) * w;
    s->ptheta += (k1.ptheta + k2.ptheta * 2.0 + k3.ptheta * 2.0 + k4.ptheta) * w;
}

// KerrGeodesic.swift:116-121.
double orc_kerr_radius_of(double a, const double point[3]) {
    double a2 = a * a;
    double rho2 = point[0] * point[0] + point[1] * point[1] + point[2] * point[2];
    double half = 0.5 * (rho2 - a2);
    return std::sqrt(half + std::sqrt(half * half + a2 * point[2] * point[2]));
}

// KerrGeodesic.swift:128-168: static observer tetrad, local energy one.
void orc_kerr_launch(double a, const double origin[3], const double direction[3], orc_kerr_state* out) {
    double r = orc_kerr_radius_of(a, origin);
    double rho = std::sqrt(origin[0] * origin[0] + origin[1] * origin[1] + origin[2] * origin[2]);
    double theta = std::acos(std::min(std::max(origin[2] / rho, -1.0), 1.0));
    double phi = std::atan2(origin[1], origin[0]);
    v3 d = orc::normalize(orc::make(direction));
    double st = std::sin(theta), ct = std::cos(theta);
    v3 r_hat = {st * std::cos(phi), st * std::sin(phi), ct};
    v3 theta_hat = {ct * std::cos(phi), ct * std::sin(phi), -st};
    v3 phi_hat = {-std::sin(phi), std::cos(phi), 0.0};
    double nr = orc::dot(d, r_hat), nth = orc::dot(d, theta_hat), nph = orc::dot(d, phi_hat);
    double delta = kdelta(a, r), sigma = ksigma(a, r, theta);
    double sin2 = std::max(st * st, 1e-12);
    double gtt = -(1.0 - 2.0 * r / sigma);
    double gtphi = -2.0 * a * r * sin2 / sigma;
    double gphiphi = (r * r + a * a + 2.0 * a * a * r * sin2 / sigma) * sin2;
    double grr = sigma / delta, gthth = sigma;
    double ut = 1.0 / std::sqrt(-gtt);
    double phi_norm = std::sqrt(gphiphi - gtphi * gtphi / gtt);
    double pt = ut + nph * (-gtphi / gtt) / phi_norm;
    double pphi = nph / phi_norm;
    out->r = r;
    out->theta = theta;
    out->phi = phi;
    out->energy = -(gtt * pt + gtphi * pphi);
    out->angular_momentum = gtphi * pt + gphiphi * pphi;
    out->pr = grr * (nr / std::sqrt(grr));
    out->ptheta = gthth * (nth / std::sqrt(gthth));
}

// KerrGeodesic.swift:170-174.
void orc_kerr_position(double a, const orc_kerr_state* s, double out[3]) {
    double radial = std::sqrt(s->r * s->r + a * a);
    double st = std::sin(s->theta);
    out[0] = radial * st * std::cos(s->phi);
    out[1] = radial * st * std::sin(s->phi);
    out[2] = s->r * std::cos(s->theta);
}

// The direction convention of the kernel (Kerr.h:158-173), which the Swift
// oracle shares but does not expose, restated in FP64.
void orc_kerr_direction(double a, const orc_kerr_state* s, double out[3]) {
    double delta = kdelta(a, s->r), sigma = ksigma(a, s->r, s->theta);
    double sin2 = std::max(std::sin(s->theta) * std::sin(s->theta), 1e-12);
    double rdot = delta * s->pr / sigma;
    double thdot = s->ptheta / sigma;
    double p, g;
    potentials(a, *s, p, g);
    double phdot = (a * p / delta + s->angular_momentum / sin2 - a * s->energy) / sigma;
    double st = std::sin(s->theta), ct = std::cos(s->theta);
    v3 r_hat = {st * std::cos(s->phi), st * std::sin(s->phi), ct};
    v3 theta_hat = {ct * std::cos(s->phi), ct * std::sin(s->phi), -st};
    v3 phi_hat = {-std::sin(s->phi), std::cos(s->phi), 0.0};
    v3 local = rdot * r_hat + s->r * thdot * theta_hat + s->r * st * phdot * phi_hat;
    double len = orc::length(local);
    orc::store(len > 0 ? local / len : r_hat, out);
}

// KerrGeodesic.swift:214-221.
double orc_kerr_step_length(double a, const orc_kerr_state* s, const orc_kerr_settings* settings) {
    double above = s->r - orc_kerr_outer_horizon(a);
    double near_h = 0.2 * above;
    double pole = std::min(std::max(std::sin(s->theta) / KERR_POLE_SINE, 0.05), 1.0);
    return std::min(std::max(std::min(settings->step_factor * s->r, near_h) * pole, 0.4 * settings->step_min),
                    settings->step_max);
}

// KerrGeodesic.swift:223-244.
void orc_kerr_integrate(double a, const orc_kerr_state* launch, const orc_kerr_settings* settings, orc_kerr_result* out) {
    orc_kerr_state s = *launch;
    double carter0 = orc_kerr_carter(a, launch);
    double scale = std::max(std::fabs(carter0), launch->energy * launch->energy);
    double drift = 0.0;
    int64_t steps = 0;
    double capture = orc_kerr_outer_horizon(a) + settings->capture_margin;
    int outcome = 2;
    while (steps < settings->step_cap) {
        orc_kerr_step(a, &s, orc_kerr_step_length(a, &s, settings));
        ++steps;
        double carter_drift = std::fabs(orc_kerr_carter(a, &s) - carter0) / scale;
        double null_drift = std::fabs(orc_kerr_null_residual(a, &s)) / (scale * std::max(s.r * s.r, 1.0));
        drift = std::max(drift, std::max(carter_drift, null_drift));
        if (s.r <= capture) {
            outcome = 0;
            break;
        }
        if (s.r >= settings->escape_radius) {
            outcome = 1;
            break;
        }
    }
    out->state = s;
    out->outcome = outcome;
    out->steps = steps;
    out->drift = drift;
}

// KerrGeodesic.swift:249-262.
double orc_kerr_bisect_critical_xi(double a, double radius, double side, const orc_kerr_settings* settings, int iterations) {
    double low = 0.0, high = 12.0 / radius;
    double origin[3] = {radius, 0.0, 0.0};
    for (int i = 0; i < iterations; ++i) {
        double aim = 0.5 * (low + high);
        double dir[3] = {-1.0, side * aim, 0.0};
        orc_kerr_state launch;
        orc_kerr_launch(a, origin, dir, &launch);
        orc_kerr_result result;
        orc_kerr_integrate(a, &launch, settings, &result);
        if (result.outcome == 0) low = aim;
        else high = aim;
    }
    double dir[3] = {-1.0, side * 0.5 * (low + high), 0.0};
    orc_kerr_state boundary;
    orc_kerr_launch(a, origin, dir, &boundary);
    return boundary.angular_momentum / boundary.energy;
}

// KerrGeodesic.swift:267-273, Bardeen 1973.
double orc_kerr_critical_xi(double a, double rc) {
    return (rc * rc * (3.0 - rc) - a * a * (rc + 1.0)) / (a * (rc - 1.0));
}

double orc_kerr_critical_eta(double a, double rc) {
    return rc * rc * rc * (4.0 * a * a - rc * (rc - 3.0) * (rc - 3.0)) / (a * a * (rc - 1.0) * (rc - 1.0));
}

void orc_kerr_photon_orbits(double a, double out[2]) {
    out[0] = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(-a)));
    out[1] = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(a)));
}

}  // extern "C"
