// FP64 Kerr null geodesics in Cartesian Kerr-Schild coordinates: the
// integrator form the reference names as its future work for the polar
// seam (docs/lab/07-kerr.md:33-45), restated here as the checker of the
// GPU's Kerr-Schild form. Test infrastructure: only tests/ load it.
//
// Coordinates (t~, x, y, z), signature -+++, M = 1:
//   g = eta + f l l,  f = 2 r^3 / (r^4 + a^2 z^2),
//   l = (1, (r x + a y) / (r^2 + a^2), (r y - a x) / (r^2 + a^2), z / r),
// with r the root of (x^2 + y^2) / (r^2 + a^2) + z^2 / r^2 = 1. They relate
// to Boyer-Lindquist (t, r, theta, phi) by t~ = t + int 2r/Delta dr,
// phi~ = phi + int a/Delta dr and x + i y = (r + i a) sin(theta) e^{i phi~},
// z = r cos(theta): the pull-back of g through that map is the
// Boyer-Lindquist metric of KerrGeodesic.swift (test_ks_metric_is_kerr).
// The metric is regular on the spin axis, where Boyer-Lindquist theta is
// not, so rays that pass over the pole are integrated without the clamped
// 1/sin^2 terms.
//
// The Hamiltonian of a photon with p_t = -E is
//   H = (1/2) (-E^2 + |p|^2 - f (E + l.p)^2),
//   dx/dlambda = p - f (E + l.p) l,
//   dp/dlambda = (1/2) (E + l.p)^2 grad f + f (E + l.p) (grad l) p,
// with the affine parameter of the Boyer-Lindquist form, so the reference's
// step lengths keep their meaning.
#include <algorithm>
#include <cmath>

#include "oracle.h"
#include "oracle_internal.hpp"

namespace {

struct Geometry {
    double r, q, d;     // r, r^2 + a^2, r^4 + a^2 z^2
    double f;           // 2 r^3 / d
    double l[3];        // spatial l
    double dr[3];       // grad r
    double fr, fz;      // df/dr at fixed z, df/dz at fixed r
    double lr[3];       // dl_k/dr at fixed x, y, z
};

double ks_radius(double a, double x, double y, double z) {
    double half = 0.5 * (x * x + y * y + z * z - a * a);
    return std::sqrt(half + std::sqrt(half * half + a * a * z * z));
}

Geometry geometry(double a, double x, double y, double z) {
    Geometry g;
    double r = ks_radius(a, x, y, z);
    g.r = r;
    g.q = r * r + a * a;
    g.d = r * r * r * r + a * a * z * z;
    g.f = 2.0 * r * r * r / g.d;
    g.l[0] = (r * x + a * y) / g.q;
    g.l[1] = (r * y - a * x) / g.q;
    g.l[2] = z / r;
    double r3 = r * r * r;
    g.dr[0] = x * r3 / g.d;
    g.dr[1] = y * r3 / g.d;
    g.dr[2] = z * r * g.q / g.d;
    g.fr = 2.0 * r * r * (3.0 * a * a * z * z - r * r * r * r) / (g.d * g.d);
    g.fz = -4.0 * a * a * z * r3 / (g.d * g.d);
    double q2 = g.q * g.q;
    g.lr[0] = (x * (a * a - r * r) - 2.0 * a * r * y) / q2;
    g.lr[1] = (y * (a * a - r * r) + 2.0 * a * r * x) / q2;
    g.lr[2] = -z / (r * r);
    return g;
}

// int a / Delta dr (the phi~ - phi shift) up to a constant.
double phi_shift(double a, double r) {
    double root = std::sqrt(1.0 - a * a);
    double rp = 1.0 + root, rm = 1.0 - root;
    return a / (rp - rm) * std::log(std::fabs((r - rp) / (r - rm)));
}

struct Rhs {
    double x[3], p[3];
};

Rhs derivative(double a, const orc_ks_state& s) {
    Geometry g = geometry(a, s.x, s.y, s.z);
    const double p[3] = {s.px, s.py, s.pz};
    double lp = g.l[0] * p[0] + g.l[1] * p[1] + g.l[2] * p[2];
    double big = s.energy + lp;
    Rhs out;
    for (int i = 0; i < 3; ++i) out.x[i] = p[i] - g.f * big * g.l[i];
    double A = g.lr[0] * p[0] + g.lr[1] * p[1] + g.lr[2] * p[2];
    const double explicit_l[3] = {(g.r * p[0] - a * p[1]) / g.q, (a * p[0] + g.r * p[1]) / g.q, p[2] / g.r};
    for (int i = 0; i < 3; ++i) {
        double grad_f = g.fr * g.dr[i] + (i == 2 ? g.fz : 0.0);
        out.p[i] = 0.5 * big * big * grad_f + g.f * big * (A * g.dr[i] + explicit_l[i]);
    }
    return out;
}

orc_ks_state axpy(const orc_ks_state& s, const Rhs& k, double h) {
    orc_ks_state o = s;
    o.x += k.x[0] * h;
    o.y += k.x[1] * h;
    o.z += k.x[2] * h;
    o.px += k.p[0] * h;
    o.py += k.p[1] * h;
    o.pz += k.p[2] * h;
    return o;
}

}  // namespace

void orc_ks_metric(double a, double x, double y, double z, double g[16]) {
    Geometry geo = geometry(a, x, y, z);
    const double l[4] = {1.0, geo.l[0], geo.l[1], geo.l[2]};
    for (int m = 0; m < 4; ++m)
        for (int n = 0; n < 4; ++n) g[4 * m + n] = (m == n ? (m == 0 ? -1.0 : 1.0) : 0.0) + geo.f * l[m] * l[n];
}

double orc_ks_radius(double a, double x, double y, double z) { return ks_radius(a, x, y, z); }

// Boyer-Lindquist covariant momenta to Kerr-Schild: raise with the
// Boyer-Lindquist inverse metric, map the contravariant vector through the
// coordinate change (phi~ = phi at the launch), lower with the Kerr-Schild
// metric.
void orc_ks_from_bl(double a, const orc_kerr_state* s, orc_ks_state* out) {
    double r = s->r, th = s->theta;
    double st = std::sin(th), ct = std::cos(th);
    double sin2 = std::max(st * st, 1e-12);
    double delta = r * r - 2.0 * r + a * a, sigma = r * r + a * a * ct * ct;
    double E = s->energy, L = s->angular_momentum;
    double g_tt = -((r * r + a * a) * (r * r + a * a) - a * a * delta * sin2) / (sigma * delta);
    double g_tphi = -2.0 * a * r / (sigma * delta);
    double g_phiphi = (delta - a * a * sin2) / (sigma * delta * sin2);
    double kt = g_tt * (-E) + g_tphi * L;
    double kphi = g_tphi * (-E) + g_phiphi * L;
    double kr = delta * s->pr / sigma;
    double kth = s->ptheta / sigma;
    double kt_ks = kt + 2.0 * r / delta * kr;
    double kphi_ks = kphi + a / delta * kr;
    double cp = std::cos(s->phi), sp = std::sin(s->phi);
    double x = (r * cp - a * sp) * st, y = (r * sp + a * cp) * st, z = r * ct;
    double k[3] = {cp * st * kr + (r * cp - a * sp) * ct * kth - y * kphi_ks,
                   sp * st * kr + (r * sp + a * cp) * ct * kth + x * kphi_ks, ct * kr - r * st * kth};
    Geometry g = geometry(a, x, y, z);
    double lk = kt_ks + g.l[0] * k[0] + g.l[1] * k[1] + g.l[2] * k[2];
    out->x = x;
    out->y = y;
    out->z = z;
    out->px = k[0] + g.f * g.l[0] * lk;
    out->py = k[1] + g.f * g.l[1] * lk;
    out->pz = k[2] + g.f * g.l[2] * lk;
    out->energy = -(-kt_ks + g.f * lk);
    out->phase = phi_shift(a, r);
}

// Kerr-Schild back to Boyer-Lindquist covariant (r, theta, phi, p_r,
// p_theta, E, L): the covector components along the Boyer-Lindquist
// coordinate directions.
void orc_ks_to_bl(double a, const orc_ks_state* s, orc_kerr_state* out) {
    double r = ks_radius(a, s->x, s->y, s->z);
    double ct = std::min(std::max(s->z / r, -1.0), 1.0);
    double th = std::acos(ct), st = std::sin(th);
    double phit = std::atan2(s->y, s->x) - std::atan2(a, r);
    double cp = std::cos(phit), sp = std::sin(phit);
    double delta = r * r - 2.0 * r + a * a;
    double pt = -s->energy;
    out->r = r;
    out->theta = th;
    out->phi = phit - (phi_shift(a, r) - s->phase);
    out->energy = s->energy;
    out->angular_momentum = s->x * s->py - s->y * s->px;
    out->pr = pt * 2.0 * r / delta + s->px * (cp * st - s->y * a / delta) + s->py * (sp * st + s->x * a / delta) +
              s->pz * ct;
    out->ptheta = s->px * (r * cp - a * sp) * ct + s->py * (r * sp + a * cp) * ct - s->pz * r * st;
}

double orc_ks_null_residual(double a, const orc_ks_state* s) {
    Geometry g = geometry(a, s->x, s->y, s->z);
    double lp = g.l[0] * s->px + g.l[1] * s->py + g.l[2] * s->pz;
    double big = s->energy + lp;
    double h = 0.5 * (-s->energy * s->energy + s->px * s->px + s->py * s->py + s->pz * s->pz - g.f * big * big);
    double sigma = g.r * g.r + a * a * s->z * s->z / (g.r * g.r);
    return 2.0 * sigma * h;
}

void orc_ks_step(double a, orc_ks_state* s, double h) {
    Rhs k1 = derivative(a, *s);
    Rhs k2 = derivative(a, axpy(*s, k1, 0.5 * h));
    Rhs k3 = derivative(a, axpy(*s, k2, 0.5 * h));
    Rhs k4 = derivative(a, axpy(*s, k3, h));
    double w = h / 6.0;
    double* xs[3] = {&s->x, &s->y, &s->z};
    double* ps[3] = {&s->px, &s->py, &s->pz};
    for (int i = 0; i < 3; ++i) {
        *xs[i] += (k1.x[i] + 2.0 * k2.x[i] + 2.0 * k3.x[i] + k4.x[i]) * w;
        *ps[i] += (k1.p[i] + 2.0 * k2.p[i] + 2.0 * k3.p[i] + k4.p[i]) * w;
    }
}

// The Boyer-Lindquist step policy without its pole factor, which the
// regular axis does not need.
double orc_ks_step_length(double a, const orc_ks_state* s, const orc_kerr_settings* settings) {
    double r = ks_radius(a, s->x, s->y, s->z);
    double near_h = 0.2 * (r - orc_kerr_outer_horizon(a));
    return std::min(std::max(std::min(settings->step_factor * r, near_h), 0.4 * settings->step_min),
                    settings->step_max);
}

// The renderer's position convention, sqrt(r^2 + a^2) sin(theta) (cos phi,
// sin phi) and r cos(theta) with the Boyer-Lindquist phi: the Kerr-Schild
// point turned about z by atan2(a, r) plus the phi~ - phi shift.
void orc_ks_position(double a, const orc_ks_state* s, double out[3]) {
    double r = ks_radius(a, s->x, s->y, s->z);
    double turn = std::atan2(a, r) + phi_shift(a, r) - s->phase;
    double c = std::cos(turn), sn = std::sin(turn);
    out[0] = c * s->x + sn * s->y;
    out[1] = -sn * s->x + c * s->y;
    out[2] = s->z;
}

// Drift gauges of the Kerr-Schild form: the null residual and the Carter
// constant as the Boyer-Lindquist form measures them (scalars, the same in
// any coordinates), the Carter constant only where |sin theta| > 0.01, since
// its formula divides by sin^2 theta, and the axial angular momentum
// x p_y - y p_x, conserved exactly by the flow and checked against
// sqrt(scale) everywhere.
double orc_ks_drift(double a, const orc_ks_state* s, double carter0, double l0, double scale) {
    orc_kerr_state bl;
    orc_ks_to_bl(a, s, &bl);
    double drift = std::fabs(orc_ks_null_residual(a, s)) / (scale * std::max(bl.r * bl.r, 1.0));
    if (std::fabs(std::sin(bl.theta)) > 0.01)
        drift = std::max(drift, std::fabs(orc_kerr_carter(a, &bl) - carter0) / scale);
    return std::max(drift, std::fabs(bl.angular_momentum - l0) / std::sqrt(scale));
}

void orc_ks_integrate(double a, const orc_ks_state* launch, const orc_kerr_settings* settings, orc_kerr_result* out) {
    orc_ks_state s = *launch;
    orc_kerr_state bl;
    orc_ks_to_bl(a, launch, &bl);
    double carter0 = orc_kerr_carter(a, &bl);
    double l0 = bl.angular_momentum;
    double scale = std::max(std::fabs(carter0), launch->energy * launch->energy);
    double drift = 0.0;
    int64_t steps = 0;
    double capture = orc_kerr_outer_horizon(a) + settings->capture_margin;
    int outcome = 2;
    while (steps < settings->step_cap) {
        orc_ks_step(a, &s, orc_ks_step_length(a, &s, settings));
        ++steps;
        drift = std::max(drift, orc_ks_drift(a, &s, carter0, l0, scale));
        bl.r = orc_ks_radius(a, s.x, s.y, s.z);
        if (bl.r <= capture) {
            outcome = 0;
            break;
        }
        if (bl.r >= settings->escape_radius) {
            outcome = 1;
            break;
        }
    }
    orc_ks_to_bl(a, &s, &out->state);
    out->outcome = outcome;
    out->steps = steps;
    out->drift = drift;
}
