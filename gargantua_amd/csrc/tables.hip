// FP64 kernels for gfx950: the Schwarzschild sweep tables the march jumps
// with (Deflection.swift, DeflectionTables.swift), a log-spaced far-field
// deflection table, the Kerr shadow boundary by parallel multi-section, and
// Schwarzschild probes integrated in double precision (the reference's FP64
// integrator, which its validation judges the GPU's probes against).
// Double precision stays on these because their stability needs it: the
// sweep integrand is singular at the turning point and the shadow edge is a
// separatrix where single precision rays change fate.
#include "device_common.hpp"
#include "launch.hpp"

namespace garg {

GARG_DEV double max_impact_parameter(double r) { return r / sqrt(1.0 - R_SCHWARZSCHILD / r); }

// Table impact parameter at coordinate t (Deflection.swift:45-48).
GARG_DEV double table_b(double low, double high, double t) {
    double back = 1.0 - t;
    return low + (high - low) * (1.0 - back * back);
}

// Periapsis: smallest root of 1 - b^2 u^2 (1 - 2u) by Newton from u = 1 / b
// (Deflection.swift:59-69).
GARG_DEV double periapsis(double b) {
    double u = 1.0 / b;
    for (int i = 0; i < 60; ++i) {
        double value = 1.0 - b * b * u * u * (1.0 - 2.0 * u);
        double slope = -b * b * (2.0 * u - 6.0 * u * u);
        double next = u - value / slope;
        if (fabs(next - u) < 1e-15) {
            u = next;
            break;
        }
        u = next;
    }
    return 1.0 / u;
}

// Midpoint-rule terms of the sweep from radius r to infinity after
// u = u_max sin^2(theta) (Deflection.swift:23-39), summed over this thread's
// share of the intervals and reduced over the workgroup.
template <int BLOCK>
GARG_DEV double block_sweep(double r, double b, uint32_t intervals, double* scratch) {
    double sum = 0.0;
    if (b > 0.0) {
        double bound = fmin(b, max_impact_parameter(r));
        double u_max = 1.0 / r;
        double h = (M_PI / 2.0) / (double)intervals;
        for (uint32_t k = threadIdx.x; k < intervals; k += BLOCK) {
            double theta = ((double)k + 0.5) * h;
            double s, c;
            sincos(theta, &s, &c);
            double u = u_max * s * s;
            double inner = 1.0 - bound * bound * u * u * (1.0 - 2.0 * u);
            if (inner > 0.0) sum += bound * 2.0 * u_max * s * c / sqrt(inner);
        }
        sum *= h;
    }
    scratch[threadIdx.x] = sum;
    __syncthreads();
    for (int stride = BLOCK / 2; stride > 0; stride >>= 1) {
        if ((int)threadIdx.x < stride) scratch[threadIdx.x] += scratch[threadIdx.x + stride];
        __syncthreads();
    }
    double total = scratch[0];
    __syncthreads();
    return total;
}

constexpr int TABLE_BLOCK = 256;

// One workgroup per table entry (DeflectionTables.swift:57-64).
__global__ __launch_bounds__(TABLE_BLOCK) void sweep_table_kernel(float* out, uint32_t n, double radius, double low,
                                                                  double high, uint32_t intervals, int kind) {
    __shared__ double scratch[TABLE_BLOCK];
    uint32_t i = blockIdx.x;
    double b = table_b(low, high, (double)i / (double)(n - 1u));
    double value;
    if (kind == 0) {
        value = block_sweep<TABLE_BLOCK>(radius, b, intervals, scratch);
    } else {
        // Escape sweep: the turn back to `radius` plus the outgoing leg
        // (Deflection.swift:77-86).
        double peri = periapsis(b);
        double s_peri = block_sweep<TABLE_BLOCK>(peri, b, intervals, scratch);
        double s_r = block_sweep<TABLE_BLOCK>(radius, b, intervals, scratch);
        value = 2.0 * (s_peri - s_r) + s_r;
    }
    if (threadIdx.x == 0) out[i] = (float)value;
}

// Far-field table: alpha(b) = 2 S(r_p, b) - pi and S(r_obs, b) at
// log-spaced b, one workgroup per entry.
__global__ __launch_bounds__(TABLE_BLOCK) void far_field_kernel(double* b_out, double* alpha, double* sweep, uint32_t n,
                                                                double r_obs, double b_lo, double b_hi,
                                                                uint32_t intervals) {
    __shared__ double scratch[TABLE_BLOCK];
    uint32_t i = blockIdx.x;
    double b = b_lo * pow(b_hi / b_lo, (double)i / (double)(n - 1u));
    double peri = periapsis(b);
    double s_peri = block_sweep<TABLE_BLOCK>(peri, b, intervals, scratch);
    double s_obs = block_sweep<TABLE_BLOCK>(r_obs, b, intervals, scratch);
    if (threadIdx.x == 0) {
        b_out[i] = b;
        alpha[i] = 2.0 * s_peri - M_PI;
        sweep[i] = s_obs;
    }
}

// ---- Kerr in double precision (KerrGeodesic.swift:52-244) ----------------

struct KState {
    double r, theta, phi, pr, ptheta;
};

struct KParams {
    double a, E, L;
};

GARG_DEV KState kd_derivative(const KParams& k, const KState& s) {
    double a = k.a, e = k.E, l = k.L;
    double r = s.r;
    double delta = r * r - 2.0 * r + a * a;
    double st, ct;
    sincos(s.theta, &st, &ct);
    double sigma = r * r + a * a * ct * ct;
    double sin2 = fmax(st * st, 1e-12);
    double p = (r * r + a * a) * e - a * l;
    double g = l * l / sin2 - 2.0 * a * e * l + a * a * e * e * sin2;
    double f = delta * s.pr * s.pr + s.ptheta * s.ptheta - p * p / delta + g;
    double hamiltonian = f / (2.0 * sigma);
    double delta_prime = 2.0 * r - 2.0;
    double fr = delta_prime * s.pr * s.pr - 2.0 * p * (2.0 * r * e) / delta + p * p * delta_prime / (delta * delta);
    double ftheta = -2.0 * l * l * ct / (sin2 * fmax(st, 1e-9)) + 2.0 * a * a * e * e * st * ct;
    KState out;
    out.r = delta * s.pr / sigma;
    out.theta = s.ptheta / sigma;
    out.phi = (a * p / delta + l / sin2 - a * e) / sigma;
    out.pr = -(fr / (2.0 * sigma) - hamiltonian * 2.0 * r / sigma);
    out.ptheta = -(ftheta / (2.0 * sigma) - hamiltonian * (-2.0 * a * a * st * ct) / sigma);
    return out;
}

GARG_DEV KState kd_axpy(const KState& s, const KState& k, double h) {
    return KState{s.r + k.r * h, s.theta + k.theta * h, s.phi + k.phi * h, s.pr + k.pr * h, s.ptheta + k.ptheta * h};
}

GARG_DEV void kd_step(const KParams& k, KState& s, double h) {
    KState k1 = kd_derivative(k, s);
    KState k2 = kd_derivative(k, kd_axpy(s, k1, 0.5 * h));
    KState k3 = kd_derivative(k, kd_axpy(s, k2, 0.5 * h));
    KState k4 = kd_derivative(k, kd_axpy(s, k3, h));
    double w = h / 6.0;
    s.r += (k1.r + k2.r * 2.0 + k3.r * 2.0 + k4.r) * w;
    s.theta += (k1.theta + k2.theta * 2.0 + k3.theta * 2.0 + k4.theta) * w;
    s.phi += (k1.phi + k2.phi * 2.0 + k3.phi * 2.0 + k4.phi) * w;
    s.pr += (k1.pr + k2.pr * 2.0 + k3.pr * 2.0 + k4.pr) * w;
    s.ptheta += (k1.ptheta + k2.ptheta * 2.0 + k3.ptheta * 2.0 + k4.ptheta) * w;
}

// Static observer launch in double precision (KerrGeodesic.swift:128-168).
GARG_DEV void kd_launch(double a, const double o[3], const double dir[3], KState& s, KParams& k) {
    double a2 = a * a;
    double rho2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    double half = 0.5 * (rho2 - a2);
    double r = sqrt(half + sqrt(half * half + a2 * o[2] * o[2]));
    double rho = sqrt(rho2);
    double theta = acos(fmin(fmax(o[2] / rho, -1.0), 1.0));
    double phi = atan2(o[1], o[0]);
    double dl = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    double d[3] = {dir[0] / dl, dir[1] / dl, dir[2] / dl};
    double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
    double nr = d[0] * st * cp + d[1] * st * sp + d[2] * ct;
    double nth = d[0] * ct * cp + d[1] * ct * sp - d[2] * st;
    double nph = -d[0] * sp + d[1] * cp;
    double delta = r * r - 2.0 * r + a2;
    double sigma = r * r + a2 * ct * ct;
    double sin2 = fmax(st * st, 1e-12);
    double gtt = -(1.0 - 2.0 * r / sigma);
    double gtphi = -2.0 * a * r * sin2 / sigma;
    double gphiphi = (r * r + a2 + 2.0 * a2 * r * sin2 / sigma) * sin2;
    double grr = sigma / delta, gthth = sigma;
    double ut = 1.0 / sqrt(-gtt);
    double phi_norm = sqrt(gphiphi - gtphi * gtphi / gtt);
    double pt = ut + nph * (-gtphi / gtt) / phi_norm;
    double pphi = nph / phi_norm;
    k.a = a;
    k.E = -(gtt * pt + gtphi * pphi);
    k.L = gtphi * pt + gphiphi * pphi;
    s = KState{r, theta, phi, grr * (nr / sqrt(grr)), gthth * (nth / sqrt(gthth))};
}

// 1 when the ray is captured under the oracle's step policy
// (KerrGeodesic.swift:214-244), 0 when it escapes or runs out of steps.
GARG_DEV int kd_captured(const KParams& k, KState s, uint32_t step_cap, double escape) {
    double rplus = 1.0 + sqrt(1.0 - k.a * k.a);
    double capture = rplus + KERR_CAPTURE_MARGIN;
    for (uint32_t n = 0; n < step_cap; ++n) {
        double near_h = 0.2 * (s.r - rplus);
        double pole = fmin(fmax(sin(s.theta) / KERR_POLE_SINE, 0.05), 1.0);
        double h = fmin(fmax(fmin(STEP_RADIUS_FACTOR * s.r, near_h) * pole, 0.4 * 0.005), STEP_MAX);
        kd_step(k, s, h);
        if (s.r <= capture) return 1;
        if (s.r >= escape) return 0;
    }
    return 0;
}

// One workgroup per screen azimuth, one lane per ray of a round. Each round
// spreads the lanes over the current bracket of aim angles, and the bracket
// shrinks to the interval where captured turns into escaped
// (Validate.swift:125-150), so each round gains log2(lanes) bits.
__global__ __launch_bounds__(64) void shadow_kernel(double* edge, double* xi, uint32_t azimuths, double spin,
                                                    double r_obs, double inclination, uint32_t rounds,
                                                    uint32_t step_cap) {
    uint32_t az = blockIdx.x;
    if (az >= azimuths) return;
    const uint32_t lanes = 64;
    double psi = 2.0 * M_PI * (double)az / (double)azimuths;
    double si = sin(inclination), ci = cos(inclination);
    double origin[3] = {r_obs * si, 0.0, r_obs * ci};
    double forward[3] = {-si, 0.0, -ci};
    // right = forward x z normalized, up = right x forward
    double right[3] = {0.0, 1.0, 0.0};
    double up[3] = {right[1] * forward[2] - right[2] * forward[1], right[2] * forward[0] - right[0] * forward[2],
                    right[0] * forward[1] - right[1] * forward[0]};
    double sp = sin(psi), cp = cos(psi);
    double screen[3] = {cp * right[0] + sp * up[0], cp * right[1] + sp * up[1], cp * right[2] + sp * up[2]};
    double escape = fmax((double)R_ESCAPE, r_obs + 10.0);
    double low = 0.0, high = 12.0 / r_obs;
    for (uint32_t round = 0; round < rounds; ++round) {
        double aim = low + (high - low) * ((double)threadIdx.x + 0.5) / (double)lanes;
        double ca = cos(aim), sa = sin(aim);
        double d[3] = {ca * forward[0] + sa * screen[0], ca * forward[1] + sa * screen[1],
                       ca * forward[2] + sa * screen[2]};
        KState s;
        KParams k;
        kd_launch(spin, origin, d, s, k);
        int captured = kd_captured(k, s, step_cap, escape);
        uint64_t escaped = __ballot(!captured);
        // First escaped lane; lanes below it were captured.
        int first = escaped ? __ffsll((unsigned long long)escaped) - 1 : (int)lanes;
        double new_low = first > 0 ? low + (high - low) * ((double)(first - 1) + 0.5) / (double)lanes : low;
        double new_high =
            first < (int)lanes ? low + (high - low) * ((double)first + 0.5) / (double)lanes : high;
        low = new_low;
        high = new_high;
    }
    if (threadIdx.x == 0) {
        double aim = 0.5 * (low + high);
        edge[az] = aim;
        if (xi) {
            double ca = cos(aim), sa = sin(aim);
            double d[3] = {ca * forward[0] + sa * screen[0], ca * forward[1] + sa * screen[1],
                           ca * forward[2] + sa * screen[2]};
            KState s;
            KParams k;
            kd_launch(spin, origin, d, s, k);
            xi[az] = k.L / k.E;
        }
    }
}

// ---- FP64 Schwarzschild probes ---------------------------------------------
// Oracle/Schwarzschild.swift's integrator, one ray per lane: the ray's plane
// spanned by the launch radius e1 and the in-plane normal e2, the state
// (v, r, phi) and its affine rates in ingoing Eddington-Finkelstein form,
// classical RK4 (:121-136), the step length with the empty-leg coarse step
// when the ray is farther from the disk slab than that step (:174-184), and
// the energy / angular momentum drift against the launch (:200-220).

struct PlaneRay64 {
    double v, r, phi, vdot, rdot, phidot;
};

GARG_DEV PlaneRay64 plane_rates(const PlaneRay64& s) {
    const double r = s.r, f = 1.0 - R_SCHWARZSCHILD / r;
    const double w2 = s.phidot * s.phidot, vv_r2 = s.vdot * s.vdot / (r * r);
    return {s.vdot,
            s.rdot,
            s.phidot,
            r * w2 - vv_r2,
            2.0 * s.rdot * s.vdot / (r * r) + f * (r * w2 - vv_r2),
            -2.0 * s.rdot * s.phidot / r};
}

GARG_DEV PlaneRay64 plane_shift(const PlaneRay64& s, const PlaneRay64& k, double h) {
    return {s.v + k.v * h, s.r + k.r * h, s.phi + k.phi * h, s.vdot + k.vdot * h, s.rdot + k.rdot * h,
            s.phidot + k.phidot * h};
}

GARG_DEV void plane_rk4(PlaneRay64& s, double h) {
    const PlaneRay64 k1 = plane_rates(s);
    const PlaneRay64 k2 = plane_rates(plane_shift(s, k1, 0.5 * h));
    const PlaneRay64 k3 = plane_rates(plane_shift(s, k2, 0.5 * h));
    const PlaneRay64 k4 = plane_rates(plane_shift(s, k3, h));
    const double w = h / 6.0;
    s.v += (k1.v + 2.0 * k2.v + 2.0 * k3.v + k4.v) * w;
    s.r += (k1.r + 2.0 * k2.r + 2.0 * k3.r + k4.r) * w;
    s.phi += (k1.phi + 2.0 * k2.phi + 2.0 * k3.phi + k4.phi) * w;
    s.vdot += (k1.vdot + 2.0 * k2.vdot + 2.0 * k3.vdot + k4.vdot) * w;
    s.rdot += (k1.rdot + 2.0 * k2.rdot + 2.0 * k3.rdot + k4.rdot) * w;
    s.phidot += (k1.phidot + 2.0 * k2.phidot + 2.0 * k3.phidot + k4.phidot) * w;
}

GARG_DEV double plane_energy(const PlaneRay64& s) { return (1.0 - R_SCHWARZSCHILD / s.r) * s.vdot - s.rdot; }
GARG_DEV double plane_angular_momentum(const PlaneRay64& s) { return s.r * s.r * s.phidot; }

// The static observer's launch (Schwarzschild.swift:76-97): unit local
// energy, the direction split into radial and in-plane tangential parts;
// e1 and e2 span the ray's plane.
GARG_DEV PlaneRay64 plane_launch(const double o[3], const double d[3], double e1[3], double e2[3]) {
    const double r0 = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    const double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double unit[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = o[k] / r0;
        unit[k] = d[k] / dl;
    }
    const double radial = unit[0] * e1[0] + unit[1] * e1[1] + unit[2] * e1[2];
    for (int k = 0; k < 3; ++k) e2[k] = unit[k] - radial * e1[k];
    double across = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
    if (across < 1e-12) {
        // A purely radial ray: any plane through it.
        double helper[3] = {fabs(e1[0]) < 0.9 ? 1.0 : 0.0, fabs(e1[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        const double along = helper[0] * e1[0] + helper[1] * e1[1] + helper[2] * e1[2];
        for (int k = 0; k < 3; ++k) e2[k] = helper[k] - along * e1[k];
        const double n = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        for (int k = 0; k < 3; ++k) e2[k] /= n;
        across = 0.0;
    } else {
        for (int k = 0; k < 3; ++k) e2[k] /= across;
    }
    const double f0 = 1.0 - R_SCHWARZSCHILD / r0;
    return PlaneRay64{0.0, r0, 0.0, (1.0 + radial) / sqrt(f0), sqrt(f0) * radial, across / r0};
}

__global__ __launch_bounds__(64) void probe_fp64_kernel(const garg_ray_probe* probes, uint32_t count,
                                                        ProbeSettings64 g, garg_ray_probe_fp64_result* out) {
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= count) return;
    double o[3], d[3];
    for (int k = 0; k < 3; ++k) {
        o[k] = probes[i].origin[k];
        d[k] = probes[i].direction[k];
    }
    double e1[3], e2[3];
    PlaneRay64 s = plane_launch(o, d, e1, e2);
    const double energy = plane_energy(s), momentum = plane_angular_momentum(s);
    const double momentum_scale = fmax(fabs(momentum), 1e-3);
    double drift = 0.0;
    uint64_t steps = 0;
    uint32_t outcome = GARG_RAY_EXHAUSTED;
    while (steps < g.step_cap) {
        double sn, cs;
        sincos(s.phi, &sn, &cs);
        const double z = s.r * (cs * e1[2] + sn * e2[2]);
        const double coarse = fmin(fmax(g.empty_step_factor * s.r, g.step_min), g.empty_step_max);
        const double h = fabs(z) > DISK_SLAB_HALF_HEIGHT + coarse ? coarse
                                                                  : fmin(fmax(g.step_factor * s.r, g.step_min), g.step_max);
        plane_rk4(s, h);
        ++steps;
        drift = fmax(drift, fmax(fabs(plane_energy(s) - energy) / fabs(energy),
                                 fabs(plane_angular_momentum(s) - momentum) / momentum_scale));
        if (s.r <= g.capture_radius) {
            outcome = GARG_RAY_CAPTURED;
            break;
        }
        if (s.r >= g.escape_radius) {
            outcome = GARG_RAY_ESCAPED;
            break;
        }
    }
    // Position in the plane, and the direction a static observer there sees
    // (Schwarzschild.swift:58-67).
    double sn, cs;
    sincos(s.phi, &sn, &cs);
    const double f = fmax(1.0 - R_SCHWARZSCHILD / s.r, 1e-6);
    double dir[3], len2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double rad = cs * e1[k] + sn * e2[k], tangent = -sn * e1[k] + cs * e2[k];
        out[i].position[k] = s.r * rad;
        dir[k] = s.rdot / sqrt(f) * rad + s.r * s.phidot * tangent;
        len2 += dir[k] * dir[k];
    }
    const double len = sqrt(len2);
    for (int k = 0; k < 3; ++k) out[i].direction[k] = len > 0.0 ? dir[k] / len : cs * e1[k] + sn * e2[k];
    out[i].drift = drift;
    out[i].steps = steps;
    out[i].outcome = outcome;
    out[i].reserved = 0u;
}

// ---- FP64 lensed render ----------------------------------------------------
// Oracle/LensedRender.swift:46-91 on the GPU, one pixel per lane: the camera
// ray through the pixel centre, the FP64 integrator above with the step
// length taken at the previous point's height, and the gas gathered in
// double precision from the resolved volume with VolumeField.sample's
// trilinear weights on texel centres and zero outside (VolumeField.swift:
// 40-63): the Doppler factor cubed blended with the gravitational one by
// density, Beer-Lambert opacity, front to back until the transmittance
// floor. The reference's beaming golden judges the GPU march's image against
// this render (Validate.swift:258-291). Contraction is off in these
// functions, so each operation rounds as in a CPU evaluation.

GARG_DEV void plane_position(const PlaneRay64& s, const double e1[3], const double e2[3], double out[3]) {
#pragma clang fp contract(off)
    const double c = cos(s.phi), sn = sin(s.phi);
    for (int k = 0; k < 3; ++k) out[k] = s.r * (c * e1[k] + sn * e2[k]);
}

// The direction a static observer at the ray's point sees (Schwarzschild.swift:58-67).
GARG_DEV void plane_direction(const PlaneRay64& s, const double e1[3], const double e2[3], double out[3]) {
#pragma clang fp contract(off)
    const double c = cos(s.phi), sn = sin(s.phi);
    const double f = fmax(1.0 - R_SCHWARZSCHILD / s.r, 1e-6);
    const double radial_rate = s.rdot / sqrt(f), tangential_rate = s.r * s.phidot;
    double local[3];
    for (int k = 0; k < 3; ++k) local[k] = radial_rate * (c * e1[k] + sn * e2[k]) + tangential_rate * (-sn * e1[k] + c * e2[k]);
    const double len = sqrt(local[0] * local[0] + local[1] * local[1] + local[2] * local[2]);
    for (int k = 0; k < 3; ++k) out[k] = len > 0.0 ? local[k] / len : c * e1[k] + sn * e2[k];
}

// VolumeField.sample in double precision over the resolved fp16 volume.
GARG_DEV void field_sample64(const Render64& a, const double p[3], double e[4], double v[3]) {
#pragma clang fp contract(off)
    const int n = (int)a.size;
    const double half[3] = {a.half_xy, a.half_xy, a.half_z};
    double frac[3];
    int base[3];
    for (int k = 0; k < 3; ++k) {
        const double unit = (p[k] + half[k]) / (2.0 * half[k]);
        const double t = unit * (double)n - 0.5;
        const double b = floor(t);
        frac[k] = t - b;
        base[k] = (int)b;
    }
    for (int c = 0; c < 4; ++c) e[c] = 0.0;
    for (int c = 0; c < 3; ++c) v[c] = 0.0;
    for (int dz = 0; dz <= 1; ++dz) {
        const double wz = dz == 0 ? 1.0 - frac[2] : frac[2];
        for (int dy = 0; dy <= 1; ++dy) {
            const double wy = dy == 0 ? 1.0 - frac[1] : frac[1];
            for (int dx = 0; dx <= 1; ++dx) {
                const double wx = dx == 0 ? 1.0 - frac[0] : frac[0];
                const double w = wx * wy * wz;
                if (w == 0.0) continue;
                const int x = base[0] + dx, y = base[1] + dy, z = base[2] + dz;
                if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) continue;
                const Voxel& vox = a.voxels[((size_t)z * (size_t)n + (size_t)y) * (size_t)n + (size_t)x];
                for (int c = 0; c < 4; ++c) e[c] += (double)vox.emission[c] * w;
                for (int c = 0; c < 3; ++c) v[c] += (double)vox.velocity[c] * w;
            }
        }
    }
}

// Schwarzschild.swift:230-240: the static observer's Doppler and
// gravitational factor for gas of coordinate speed |gas|, capped at 0.85 c.
GARG_DEV double redshift_factor64(double f, const double direction[3], const double gas[3]) {
#pragma clang fp contract(off)
    const double magnitude = sqrt(gas[0] * gas[0] + gas[1] * gas[1] + gas[2] * gas[2]);
    const double speed = fmin(magnitude, 8.0);
    const double beta = speed > 0.0 ? GAS_SPEED_CEILING * tanh(speed / GAS_SPEED_CEILING) : 0.0;
    double along = 0.0;
    for (int k = 0; k < 3; ++k) along += direction[k] * (magnitude > 0.0 ? gas[k] * (beta / magnitude) : 0.0);
    const double gamma = 1.0 / sqrt(1.0 - beta * beta);
    return sqrt(fmax(f, 0.0)) / (gamma * (1.0 + along));
}

__global__ __launch_bounds__(64) void render_fp64_kernel(Render64 a) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= a.width * a.height) return;
    const uint32_t px = i % a.width, py = i / a.width;
    // The pixel's camera ray (LensedRender.swift:23-28).
    const double x = (2.0 * ((double)px + 0.5) / (double)a.width - 1.0) * a.tan_half_fov[0];
    const double y = (1.0 - 2.0 * ((double)py + 0.5) / (double)a.height) * a.tan_half_fov[1];
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = a.forward[k] + x * a.right[k] + y * a.up[k];
    const double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; ++k) d[k] = d[k] / dl;
    double e1[3], e2[3];
    PlaneRay64 s = plane_launch(a.position, d, e1, e2);
    double prev[3], color[3] = {0.0, 0.0, 0.0}, transmittance = 1.0;
    plane_position(s, e1, e2, prev);
    const ProbeSettings64& g = a.settings;
    for (uint64_t k = 0; k < g.step_cap; ++k) {
        const double coarse = fmin(fmax(g.empty_step_factor * s.r, g.step_min), g.empty_step_max);
        const double h = fabs(prev[2]) > DISK_SLAB_HALF_HEIGHT + coarse
                             ? coarse
                             : fmin(fmax(g.step_factor * s.r, g.step_min), g.step_max);
        plane_rk4(s, h);
        double pos[3];
        plane_position(s, e1, e2, pos);
        if (s.r <= g.capture_radius) break;
        if (fabs(pos[2]) < DISK_SLAB_HALF_HEIGHT && fabs(pos[0]) < a.half_xy && fabs(pos[1]) < a.half_xy &&
            transmittance > TRANSMITTANCE_FLOOR) {
            double e[4], v[3];
            field_sample64(a, pos, e, v);
            if (e[3] > 0.0) {
                const double dx = pos[0] - prev[0], dy = pos[1] - prev[1], dz = pos[2] - prev[2];
                const double path = sqrt(dx * dx + dy * dy + dz * dz);
                double g3 = 1.0;
                if (a.redshift) {
                    double dir[3];
                    plane_direction(s, e1, e2, dir);
                    const double f = 1.0 - R_SCHWARZSCHILD / s.r;
                    const double gr = redshift_factor64(f, dir, v);
                    const double gravity = sqrt(fmax(f, 0.0));
                    const double bulk = fmin(e[3] / BEAMING_DENSITY_FLOOR, 1.0);
                    g3 = gravity * gravity * gravity * (1.0 - bulk) + gr * gr * gr * bulk;
                }
                const double alpha = 1.0 - exp(-a.opacity * e[3] * path);
                const double weight = transmittance * g3 * path;
                for (int c = 0; c < 3; ++c) color[c] = color[c] + weight * e[c];
                transmittance *= 1.0 - alpha;
            }
        }
        for (int c = 0; c < 3; ++c) prev[c] = pos[c];
        if (s.r >= g.escape_radius) break;
    }
    for (int c = 0; c < 3; ++c) a.rgb[3 * (size_t)i + c] = color[c];
}

hipError_t launch_render_fp64(const Render64& a, hipStream_t stream) {
    render_fp64_kernel<<<dim3((a.width * a.height + 63u) / 64u), 64, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_probe_fp64(const garg_ray_probe* probes, uint32_t count, const ProbeSettings64& g,
                             garg_ray_probe_fp64_result* out, hipStream_t stream) {
    probe_fp64_kernel<<<dim3((count + 63u) / 64u), 64, 0, stream>>>(probes, count, g, out);
    return hipGetLastError();
}

hipError_t launch_sweep_table(float* out, uint32_t n, double radius, double low, double high, uint32_t intervals,
                              int kind, hipStream_t stream) {
    sweep_table_kernel<<<dim3(n), TABLE_BLOCK, 0, stream>>>(out, n, radius, low, high, intervals, kind);
    return hipGetLastError();
}

hipError_t launch_far_field(double* b, double* alpha, double* sweep, uint32_t n, double r_obs, double b_lo, double b_hi,
                            uint32_t intervals, hipStream_t stream) {
    far_field_kernel<<<dim3(n), TABLE_BLOCK, 0, stream>>>(b, alpha, sweep, n, r_obs, b_lo, b_hi, intervals);
    return hipGetLastError();
}

hipError_t launch_shadow(double* edge, double* xi, uint32_t azimuths, double spin, double r_obs, double inclination,
                         uint32_t rounds, uint32_t rays_per_round, uint32_t step_cap, hipStream_t stream) {
    (void)rays_per_round;  // one wave64 per azimuth: 64 rays per round
    shadow_kernel<<<dim3(azimuths), 64, 0, stream>>>(edge, xi, azimuths, spin, r_obs, inclination, rounds, step_cap);
    return hipGetLastError();
}

}  // namespace garg
