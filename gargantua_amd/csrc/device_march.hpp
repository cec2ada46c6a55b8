// Per-ray geodesic march for gfx950: Schwarzschild in ingoing
// Eddington-Finkelstein coordinates with the sweep-table jump, and Kerr in
// Boyer-Lindquist coordinates in two forms (the reference's formulas, and a
// fast form of this library). Each march is a small resumable state machine,
// so the one-launch-per-frame kernel runs it to the end in one call and the
// persistent kernel runs it a compaction period at a time.
#pragma once

#include "device_common.hpp"
#include "fast_math.hpp"

namespace garg {

// A per-step hook that does nothing.
struct NoTick {
    GARG_DEV void operator()(uint32_t) const {}
};


// ---------------------------------------------------------------------------
// Schwarzschild (Geodesic.h:9-188). The Eddington-Finkelstein time v is never
// read by anything the march outputs, so it is not carried.

struct RayState {
    float r, phi, vdot, rdot, phidot;
};

struct PlaneRay {
    RayState s;
    f3 e1, e2;
    float energy, angular_momentum;
};

GARG_DEV float schw_f(float r) { return 1.0f - GARG_F(R_SCHWARZSCHILD) / r; }
GARG_DEV float ray_energy(const RayState& s) { return schw_f(s.r) * s.vdot - s.rdot; }
GARG_DEV float ray_angular_momentum(const RayState& s) { return s.r * s.r * s.phidot; }

// Euler-Lagrange equations of L = (1/2)(-f vdot^2 + 2 vdot rdot + r^2 phidot^2).
GARG_DEV RayState ray_derivative(const RayState& s) {
    float r = s.r;
    float r2 = r * r;
    float vdd = -s.vdot * s.vdot / r2 + r * s.phidot * s.phidot;
    float rdd = 2.0f * s.rdot * s.vdot / r2 + schw_f(r) * (r * s.phidot * s.phidot - s.vdot * s.vdot / r2);
    float pdd = -2.0f * s.rdot * s.phidot / r;
    return RayState{s.rdot, s.phidot, vdd, rdd, pdd};
}

// The same right-hand sides with one hardware reciprocal of r in place of
// the four IEEE divisions (the fast form): -vdot^2/r^2 + r phidot^2 is the
// vdot equation and, times f, the bracket of the rdot equation.
GARG_DEV RayState ray_derivative_fast(const RayState& s) {
    float r = s.r;
    float ir = __builtin_amdgcn_rcpf(r);
    float ir2 = ir * ir;
    float vdd = r * s.phidot * s.phidot - s.vdot * s.vdot * ir2;
    float f = 1.0f - GARG_F(R_SCHWARZSCHILD) * ir;
    float rdd = 2.0f * s.rdot * s.vdot * ir2 + f * vdd;
    float pdd = -2.0f * s.rdot * s.phidot * ir;
    return RayState{s.rdot, s.phidot, vdd, rdd, pdd};
}

GARG_DEV RayState ray_axpy(const RayState& s, const RayState& k, float h) {
    return RayState{s.r + k.r * h, s.phi + k.phi * h, s.vdot + k.vdot * h, s.rdot + k.rdot * h, s.phidot + k.phidot * h};
}

// Classical RK4 increment (h / 6)((k1 + 2 k2) + (2 k3 + k4)); FAST takes
// the fast right-hand sides and h times 1/6.
template <bool FAST = false>
GARG_DEV RayState ray_increment(const RayState& s, float h) {
    auto derivative = [](const RayState& x) { return FAST ? ray_derivative_fast(x) : ray_derivative(x); };
    RayState k1 = derivative(s);
    RayState k2 = derivative(ray_axpy(s, k1, 0.5f * h));
    RayState k3 = derivative(ray_axpy(s, k2, 0.5f * h));
    RayState k4 = derivative(ray_axpy(s, k3, h));
    float w = FAST ? h * (1.0f / 6.0f) : h / 6.0f;
    return RayState{((k1.r + k2.r * 2.0f) + (k3.r * 2.0f + k4.r)) * w,
                    ((k1.phi + k2.phi * 2.0f) + (k3.phi * 2.0f + k4.phi)) * w,
                    ((k1.vdot + k2.vdot * 2.0f) + (k3.vdot * 2.0f + k4.vdot)) * w,
                    ((k1.rdot + k2.rdot * 2.0f) + (k3.rdot * 2.0f + k4.rdot)) * w,
                    ((k1.phidot + k2.phidot * 2.0f) + (k3.phidot * 2.0f + k4.phidot)) * w};
}

// Knuth's TwoSum: s + y exactly as t + err; err is carried into the next add.
GARG_DEV void two_sum(float& s, float& carry, float increment) {
    float y = increment + carry;
    float t = s + y;
    float bp = t - s;
    carry = (s - (t - bp)) + (y - bp);
    s = t;
}

// Dekker's Fast2Sum of s and y = sum * w + carry (one rounding): t + err
// equals s + y exactly when |s| >= |y|, which holds for r, phi and theta
// away from the pole and for p_r and p_theta away from their turning
// points; in the few steps where it does not, t is still the rounded sum
// and only that step's carry is approximate.
GARG_DEV void fast_two_sum(float& s, float& carry, float sum, float w) {
    float y = fmaf(sum, w, carry);
    float t = s + y;
    carry = y - (t - s);
    s = t;
}

GARG_DEV void ray_apply(RayState& s, RayState& carry, const RayState& inc, bool compensated) {
    if (compensated) {
        two_sum(s.r, carry.r, inc.r);
        two_sum(s.phi, carry.phi, inc.phi);
        two_sum(s.vdot, carry.vdot, inc.vdot);
        two_sum(s.rdot, carry.rdot, inc.rdot);
        two_sum(s.phidot, carry.phidot, inc.phidot);
    } else {
        s.r += inc.r;
        s.phi += inc.phi;
        s.vdot += inc.vdot;
        s.rdot += inc.rdot;
        s.phidot += inc.phidot;
    }
}

// Static observer launch with local photon energy one (Geodesic.h:63-87).
GARG_DEV PlaneRay launch_ray(f3 origin, f3 direction) {
    float r0 = length(origin);
    f3 e1 = origin / r0;
    f3 unit = normalize(direction);
    float dr = dot(unit, e1);
    f3 perpendicular = unit - e1 * dr;
    float dphi = length(perpendicular);
    f3 e2;
    if (dphi < 1e-6f) {
        f3 helper = fabsf(e1.x) < 0.9f ? f3{1.0f, 0.0f, 0.0f} : f3{0.0f, 1.0f, 0.0f};
        e2 = normalize(helper - e1 * dot(helper, e1));
        dphi = 0.0f;
    } else {
        e2 = perpendicular / dphi;
    }
    float root = sqrtf(schw_f(r0));
    PlaneRay ray;
    ray.s = RayState{r0, 0.0f, (1.0f + dr) / root, root * dr, dphi / r0};
    ray.e1 = e1;
    ray.e2 = e2;
    ray.energy = ray_energy(ray.s);
    ray.angular_momentum = ray_angular_momentum(ray.s);
    return ray;
}

GARG_DEV f3 ray_position(const PlaneRay& ray) {
    float c, s;
    sincosf(ray.s.phi, &s, &c);
    return (ray.e1 * c + ray.e2 * s) * ray.s.r;
}

// Direction measured by a static observer at the ray's radius (Geodesic.h:94-100).
GARG_DEV f3 ray_direction(const PlaneRay& ray) {
    float c, s;
    sincosf(ray.s.phi, &s, &c);
    f3 radial = ray.e1 * c + ray.e2 * s;
    f3 tangential = ray.e2 * c - ray.e1 * s;
    f3 local = radial * (ray.s.rdot / sqrtf(fmaxf(schw_f(ray.s.r), 1e-6f))) + tangential * (ray.s.r * ray.s.phidot);
    float len = length(local);
    return len > 0.0f ? local / len : radial;
}

// The fast form's position and direction: fast_sincos and hardware
// reciprocal square roots.
GARG_DEV f3 ray_position_fast(const PlaneRay& ray) {
    float c, s;
    fast_sincos(ray.s.phi, &s, &c);
    return (ray.e1 * c + ray.e2 * s) * ray.s.r;
}

GARG_DEV f3 ray_direction_fast(const PlaneRay& ray) {
    float c, s;
    fast_sincos(ray.s.phi, &s, &c);
    f3 radial = ray.e1 * c + ray.e2 * s;
    f3 tangential = ray.e2 * c - ray.e1 * s;
    float f = fmaxf(1.0f - GARG_F(R_SCHWARZSCHILD) * __builtin_amdgcn_rcpf(ray.s.r), 1e-6f);
    f3 local = radial * (ray.s.rdot * __builtin_amdgcn_rsqf(f)) + tangential * (ray.s.r * ray.s.phidot);
    float len2 = dot(local, local);
    return len2 > 0.0f ? local * __builtin_amdgcn_rsqf(len2) : radial;
}

// Table coordinate of b: tables sample b = low + (high - low)(1 - (1 - t)^2)
// at t = i / (N - 1), remapped to texel centres (Geodesic.h:106-110).
GARG_DEV float sweep_coordinate(float low, float high, float b) {
    float x = fminf(fmaxf((b - low) / (high - low), 0.0f), 1.0f);
    float t = 1.0f - sqrtf(1.0f - x);
    return (t * ((float)SWEEP_TABLE_SIZE - 1.0f) + 0.5f) / (float)SWEEP_TABLE_SIZE;
}

GARG_DEV float sweep_lookup(const float* table, float low, float high, float b) {
    return table_linear(table, (int)SWEEP_TABLE_SIZE, sweep_coordinate(low, high, b));
}

GARG_DEV f3 asymptotic_direction(const PlaneRay& ray, float phi) {
    float c, s;
    sincosf(phi, &s, &c);
    return ray.e1 * c + ray.e2 * s;
}

// Step length: coarse on empty legs farther from the slab than that step
// (Geodesic.h:177-181).
GARG_DEV float ray_step_length(float z, float r, const garg_geodesic_settings& g) {
    float coarse = fminf(fmaxf(g.empty_step_factor * r, g.step_min), g.empty_step_max);
    if (fabsf(z) > GARG_F(DISK_SLAB_HALF_HEIGHT) + coarse) return coarse;
    return fminf(fmaxf(g.step_factor * r, g.step_min), g.step_max);
}

// Largest relative deviation of E or L from the launch values (Geodesic.h:184-188).
GARG_DEV float ray_drift(const PlaneRay& ray) {
    float e = fabsf(ray_energy(ray.s) - ray.energy) / fabsf(ray.energy);
    float l = fabsf(ray_angular_momentum(ray.s) - ray.angular_momentum) / fmaxf(fabsf(ray.angular_momentum), 1e-3f);
    return fmaxf(e, l);
}

struct Prepared {
    PlaneRay ray;
    bool sky_only;
    f3 sky;
};

// Skip the empty leg from the camera to the integration sphere with the
// sweep tables (Geodesic.h:130-163).
GARG_DEV Prepared prepare_ray(f3 origin, f3 direction, const garg_march_uniforms& u, const Scene& sc) {
    Prepared p;
    p.ray = launch_ray(origin, direction);
    p.sky_only = false;
    p.sky = f3{0, 0, 0};
    if (u.sphere_radius <= 0.0f || p.ray.s.r <= u.sphere_radius) return p;
    float b = fabsf(p.ray.angular_momentum) / p.ray.energy;
    if (p.ray.s.rdot >= 0.0f) {
        float sweep = sweep_lookup(sc.camera_table, 0.0f, u.camera_max_b, b);
        p.sky_only = true;
        p.sky = asymptotic_direction(p.ray, p.ray.s.phi + sweep);
        return p;
    }
    if (b >= u.sphere_max_b) {
        float sweep = sweep_lookup(sc.sky_table, u.sky_min_b, u.sky_max_b, b);
        p.sky_only = true;
        p.sky = asymptotic_direction(p.ray, p.ray.s.phi + sweep);
        return p;
    }
    float sweep = sweep_lookup(sc.sphere_table, 0.0f, u.sphere_table_max_b, b) -
                  sweep_lookup(sc.camera_table, 0.0f, u.camera_max_b, b);
    float r = u.sphere_radius;
    float f = schw_f(r);
    float E = p.ray.energy, L = p.ray.angular_momentum;
    float phidot = L / (r * r);
    float rdot = -sqrtf(fmaxf(E * E - f * L * L / (r * r), 0.0f));
    float vdot = (E + rdot) / f;
    p.ray.s = RayState{r, p.ray.s.phi + sweep, vdot, rdot, phidot};
    return p;
}

GARG_DEV f3 sphere_exit_direction(const PlaneRay& ray, const garg_march_uniforms& u, const Scene& sc) {
    float b = fabsf(ray.angular_momentum) / ray.energy;
    float sweep = sweep_lookup(sc.sphere_table, 0.0f, u.sphere_table_max_b, b);
    return asymptotic_direction(ray, ray.s.phi + sweep);
}

// The Schwarzschild loop of march.metal:51-117 as a resumable machine. The
// reference form (FAST false) evaluates Geodesic.h as written; the fast form
// takes one hardware reciprocal of r per right-hand side, fast_sincos for
// the position and hardware square roots, on the same steps.
template <bool FAST>
struct SchwMarchT {
    PlaneRay ray;
    RayState carry;
    Gather g;
    f3 prev;
    uint32_t steps;
    uint32_t outcome;
    bool left_sphere;
    bool done;

    GARG_DEV void begin(const PlaneRay& launched) {
        ray = launched;
        carry = RayState{0, 0, 0, 0, 0};
        g.reset();
        prev = position();
        steps = 0;
        outcome = GARG_RAY_EXHAUSTED;
        left_sphere = false;
        done = false;
    }

    // Advance by at most `budget` steps; `gather` false skips the volume.
    // `tick` runs once per step of the wave (scalar work only).
    template <class Tick = NoTick>
    GARG_DEV void run(const garg_march_uniforms& u, const Scene& sc, bool gather, bool compensated, uint32_t budget,
                      Tick tick = Tick{}) {
        const garg_geodesic_settings& gs = u.geodesic;
        const bool has_volume = gather && sc.volume.voxels != nullptr;
        for (uint32_t n = 0; n < budget && !done; ++n) {
            tick(n);
            if (steps >= gs.step_cap) {
                done = true;
                break;
            }
            float h = ray_step_length(prev.z, ray.s.r, gs);
            RayState before = ray.s, carry_before = carry;
            ray_apply(ray.s, carry, ray_increment<FAST>(ray.s, h), compensated);
            steps += 1;
            if (u.sphere_radius > 0.0f && ray.s.r >= u.sphere_radius && ray.s.rdot > 0.0f && before.r < u.sphere_radius) {
                // Land on the sphere: the sweep table is exact only from there.
                float hh = h * (u.sphere_radius - before.r) / (ray.s.r - before.r);
                ray.s = before;
                carry = carry_before;
                ray_apply(ray.s, carry, ray_increment<FAST>(before, hh), compensated);
            }
            f3 p = position();
            f3 dp = p - prev;
            float path = FAST ? __builtin_amdgcn_sqrtf(dot(dp, dp)) : length(dp);
            g.travelled += path;
            if (ray.s.r <= gs.capture_radius) {
                outcome = GARG_RAY_CAPTURED;
                done = true;
                break;
            }
            if (has_volume) {
                const PlaneRay& rr = ray;
                g.sample<FAST>(sc.volume, u.opacity_scale, u.redshift != 0u, p, path, schw_f(ray.s.r),
                         [&]() { return FAST ? ray_direction_fast(rr) : ray_direction(rr); });
            }
            prev = p;
            if (u.sphere_radius > 0.0f && ray.s.r >= u.sphere_radius && ray.s.rdot > 0.0f) {
                outcome = GARG_RAY_ESCAPED;
                left_sphere = true;
                done = true;
                break;
            }
            if (ray.s.r >= gs.escape_radius) {
                outcome = GARG_RAY_ESCAPED;
                done = true;
                break;
            }
        }
        if (!done && steps >= gs.step_cap) done = true;
    }

    // Whether the ray falls into the hole (march.metal:44-47).
    GARG_DEV bool falls_in() const {
        return outcome == GARG_RAY_CAPTURED ||
               (outcome == GARG_RAY_EXHAUSTED && ray.s.rdot < 0.0f &&
                fabsf(ray.angular_momentum) < GARG_F(B_CRITICAL) * ray.energy);
    }

    GARG_DEV f3 sky_direction(const garg_march_uniforms& u, const Scene& sc) const {
        return left_sphere ? sphere_exit_direction(ray, u, sc) : direction();
    }

    GARG_DEV f3 position() const { return FAST ? ray_position_fast(ray) : ray_position(ray); }
    GARG_DEV f3 direction() const { return FAST ? ray_direction_fast(ray) : ray_direction(ray); }
};

// ---------------------------------------------------------------------------
// Kerr (Kerr.h:11-186). E and L are parameters of the separable Hamiltonian;
// the Carter constant and the null residual are the drift gauges.

struct KerrRayState {
    float r, theta, phi, pr, ptheta;
    float energy, angular_momentum;
};

GARG_DEV float kerr_outer_horizon(float a) { return 1.0f + sqrtf(fmaxf(1.0f - a * a, 0.0f)); }
GARG_DEV float kerr_delta(float a, float r) { return r * r - 2.0f * r + a * a; }
GARG_DEV float kerr_sigma(float a, float r, float theta) {
    float c = cosf(theta);
    return r * r + a * a * c * c;
}

// Boyer-Lindquist radius of a Cartesian point (Kerr.h:101-106).
GARG_DEV float kerr_radius_of(float a, f3 p) {
    float a2 = a * a;
    float rho2 = dot(p, p);
    float halved = 0.5f * (rho2 - a2);
    return sqrtf(halved + sqrtf(halved * halved + a2 * p.z * p.z));
}

// Launch from a static observer tetrad along a locally measured direction
// decomposed against the spherical frame (Kerr.h:110-148).
GARG_DEV KerrRayState kerr_launch(float a, f3 origin, f3 direction) {
    float r = kerr_radius_of(a, origin);
    float rho = length(origin);
    float theta = acosf(fminf(fmaxf(origin.z / rho, -1.0f), 1.0f));
    float phi = atan2f(origin.y, origin.x);
    f3 d = normalize(direction);
    float st, ct, sp, cp;
    sincosf(theta, &st, &ct);
    sincosf(phi, &sp, &cp);
    f3 r_hat{st * cp, st * sp, ct};
    f3 theta_hat{ct * cp, ct * sp, -st};
    f3 phi_hat{-sp, cp, 0.0f};
    float nr = dot(d, r_hat), nth = dot(d, theta_hat), nph = dot(d, phi_hat);
    float delta = kerr_delta(a, r);
    float sigma = r * r + a * a * ct * ct;
    float sin2 = fmaxf(st * st, 1e-8f);
    float gtt = -(1.0f - 2.0f * r / sigma);
    float gtphi = -2.0f * a * r * sin2 / sigma;
    float gphiphi = (r * r + a * a + 2.0f * a * a * r * sin2 / sigma) * sin2;
    float grr = sigma / delta;
    float gthth = sigma;
    float ut = 1.0f / sqrtf(-gtt);
    float phi_norm = sqrtf(gphiphi - gtphi * gtphi / gtt);
    float pt = ut + nph * (-gtphi / gtt) / phi_norm;
    float pphi = nph / phi_norm;
    KerrRayState s;
    s.r = r;
    s.theta = theta;
    s.phi = phi;
    s.energy = -(gtt * pt + gtphi * pphi);
    s.angular_momentum = gtphi * pt + gphiphi * pphi;
    s.pr = grr * (nr / sqrtf(grr));
    s.ptheta = gthth * (nth / sqrtf(gthth));
    return s;
}

GARG_DEV f3 kerr_position_sc(float a, float r, float st, float ct, float phi) {
    float sp, cp;
    sincosf(phi, &sp, &cp);
    float radial = sqrtf(r * r + a * a);
    return f3{radial * st * cp, radial * st * sp, r * ct};
}

// Direction of travel against the spherical frame from the coordinate
// velocities (Kerr.h:158-173).
GARG_DEV f3 kerr_direction_from(float r, float st, float ct, float phi, float rdot, float thdot, float phdot) {
    float sp, cp;
    sincosf(phi, &sp, &cp);
    f3 r_hat{st * cp, st * sp, ct};
    f3 theta_hat{ct * cp, ct * sp, -st};
    f3 phi_hat{-sp, cp, 0.0f};
    f3 local = r_hat * rdot + theta_hat * (r * thdot) + phi_hat * (r * st * phdot);
    float len = length(local);
    return len > 0.0f ? local / len : r_hat;
}

// --- reference form: the formulas of Kerr.h as written ---------------------
namespace kref {

GARG_DEV float2 potentials(float a, const KerrRayState& s) {
    float p = (s.r * s.r + a * a) * s.energy - a * s.angular_momentum;
    float sin2 = fmaxf(sinf(s.theta) * sinf(s.theta), 1e-8f);
    float g = s.angular_momentum * s.angular_momentum / sin2 - 2.0f * a * s.energy * s.angular_momentum +
              a * a * s.energy * s.energy * sin2;
    return make_float2(p, g);
}

GARG_DEV float null_residual(float a, const KerrRayState& s) {
    float delta = kerr_delta(a, s.r);
    float2 pg = potentials(a, s);
    return delta * s.pr * s.pr + s.ptheta * s.ptheta - pg.x * pg.x / delta + pg.y;
}

GARG_DEV float carter(float a, const KerrRayState& s) {
    float c = cosf(s.theta);
    float sin2 = fmaxf(sinf(s.theta) * sinf(s.theta), 1e-8f);
    float term = s.angular_momentum * s.angular_momentum / sin2 - a * a * s.energy * s.energy;
    return s.ptheta * s.ptheta + c * c * term;
}

// Hamilton's equations including the off-shell Sigma terms (Kerr.h:62-99).
GARG_DEV KerrRayState derivative(float a, const KerrRayState& s) {
    float r = s.r, theta = s.theta;
    float e = s.energy, l = s.angular_momentum;
    float delta = kerr_delta(a, r);
    float sigma = kerr_sigma(a, r, theta);
    float delta_prime = 2.0f * r - 2.0f;
    float st = sinf(theta), ct = cosf(theta);
    float sin2 = fmaxf(st * st, 1e-8f);
    float2 pg = potentials(a, s);
    float p = pg.x, g = pg.y;
    float f = delta * s.pr * s.pr + s.ptheta * s.ptheta - p * p / delta + g;
    float hamiltonian = f / (2.0f * sigma);
    float p_prime = 2.0f * r * e;
    float fr = delta_prime * s.pr * s.pr - 2.0f * p * p_prime / delta + p * p * delta_prime / (delta * delta);
    float ftheta = -2.0f * l * l * ct / (sin2 * fmaxf(st, 1e-4f)) + 2.0f * a * a * e * e * st * ct;
    float sigma_r = 2.0f * r;
    float sigma_theta = -2.0f * a * a * st * ct;
    KerrRayState out = s;
    out.r = delta * s.pr / sigma;
    out.theta = s.ptheta / sigma;
    out.phi = (a * p / delta + l / sin2 - a * e) / sigma;
    out.pr = -(fr / (2.0f * sigma) - hamiltonian * sigma_r / sigma);
    out.ptheta = -(ftheta / (2.0f * sigma) - hamiltonian * sigma_theta / sigma);
    return out;
}

GARG_DEV KerrRayState axpy(KerrRayState s, const KerrRayState& k, float h) {
    s.r += k.r * h;
    s.theta += k.theta * h;
    s.phi += k.phi * h;
    s.pr += k.pr * h;
    s.ptheta += k.ptheta * h;
    return s;
}

GARG_DEV KerrRayState step(float a, const KerrRayState& s, float h) {
    KerrRayState k1 = derivative(a, s);
    KerrRayState k2 = derivative(a, axpy(s, k1, 0.5f * h));
    KerrRayState k3 = derivative(a, axpy(s, k2, 0.5f * h));
    KerrRayState k4 = derivative(a, axpy(s, k3, h));
    KerrRayState out = s;
    float w = h / 6.0f;
    out.r += ((k1.r + k2.r * 2.0f) + (k3.r * 2.0f + k4.r)) * w;
    out.theta += ((k1.theta + k2.theta * 2.0f) + (k3.theta * 2.0f + k4.theta)) * w;
    out.phi += ((k1.phi + k2.phi * 2.0f) + (k3.phi * 2.0f + k4.phi)) * w;
    out.pr += ((k1.pr + k2.pr * 2.0f) + (k3.pr * 2.0f + k4.pr)) * w;
    out.ptheta += ((k1.ptheta + k2.ptheta * 2.0f) + (k3.ptheta * 2.0f + k4.ptheta)) * w;
    return out;
}

GARG_DEV f3 position(float a, const KerrRayState& s) {
    float radial = sqrtf(s.r * s.r + a * a);
    float st = sinf(s.theta);
    return f3{radial * st * cosf(s.phi), radial * st * sinf(s.phi), s.r * cosf(s.theta)};
}

GARG_DEV f3 direction(float a, const KerrRayState& s) {
    float delta = kerr_delta(a, s.r);
    float sigma = kerr_sigma(a, s.r, s.theta);
    float sin2 = fmaxf(sinf(s.theta) * sinf(s.theta), 1e-8f);
    float rdot = delta * s.pr / sigma;
    float thdot = s.ptheta / sigma;
    float2 pg = potentials(a, s);
    float phdot = (a * pg.x / delta + s.angular_momentum / sin2 - a * s.energy) / sigma;
    return kerr_direction_from(s.r, sinf(s.theta), cosf(s.theta), s.phi, rdot, thdot, phdot);
}

// Kerr.h:177-186.
GARG_DEV float step_length(float a, const KerrRayState& s, const garg_geodesic_settings& g) {
    float near_h = 0.2f * (s.r - kerr_outer_horizon(a));
    float pole = fminf(fmaxf(sinf(s.theta) / GARG_F(KERR_POLE_SINE), 0.05f), 1.0f);
    return fminf(fmaxf(fminf(g.step_factor * s.r, near_h) * pole, GARG_F(KERR_STEP_FLOOR)), g.step_max);
}

}  // namespace kref

// The Kerr loop of march.metal:123-184 in the reference form.
struct KerrRefMarch {
    KerrRayState s;
    float a, carter0, scale, capture;
    Gather g;
    f3 prev;
    float drift;
    uint32_t steps;
    uint32_t outcome;
    bool done;

    GARG_DEV void begin(float spin, const KerrRayState& launched) {
        a = spin;
        s = launched;
        carter0 = kref::carter(a, s);
        scale = fmaxf(fabsf(carter0), s.energy * s.energy);
        capture = kerr_outer_horizon(a) + GARG_F(KERR_CAPTURE_MARGIN);
        g.reset();
        prev = kref::position(a, s);
        drift = 0.0f;
        steps = 0;
        outcome = GARG_RAY_EXHAUSTED;
        done = false;
    }

    GARG_DEV void run(const garg_march_uniforms& u, const Scene& sc, bool gather, uint32_t budget) {
        const bool has_volume = gather && sc.volume.voxels != nullptr;
        for (uint32_t n = 0; n < budget && !done; ++n) {
            if (steps >= u.geodesic.step_cap) {
                done = true;
                break;
            }
            s = kref::step(a, s, kref::step_length(a, s, u.geodesic));
            steps += 1;
            f3 p = kref::position(a, s);
            float path = length(p - prev);
            g.travelled += path;
            float carter_drift = fabsf(kref::carter(a, s) - carter0) / scale;
            float null_drift = fabsf(kref::null_residual(a, s)) / (scale * fmaxf(s.r * s.r, 1.0f));
            drift = fmaxf(drift, fmaxf(carter_drift, null_drift));
            if (s.r <= capture) {
                outcome = GARG_RAY_CAPTURED;
                done = true;
                break;
            }
            if (has_volume) {
                float f = 1.0f - 2.0f * s.r / kerr_sigma(a, s.r, s.theta);
                const KerrRayState& ss = s;
                float aa = a;
                g.sample(sc.volume, u.opacity_scale, u.redshift != 0u, p, path, fmaxf(f, 0.0f),
                         [&]() { return kref::direction(aa, ss); });
            }
            prev = p;
            if (s.r >= u.geodesic.escape_radius) {
                outcome = GARG_RAY_ESCAPED;
                done = true;
                break;
            }
        }
        if (!done && steps >= u.geodesic.step_cap) done = true;
    }

    GARG_DEV bool falls_in() const {
        return outcome == GARG_RAY_CAPTURED || (outcome == GARG_RAY_EXHAUSTED && s.pr < 0.0f);
    }
    GARG_DEV f3 final_direction() const { return kref::direction(a, s); }
    GARG_DEV f3 final_position() const { return kref::position(a, s); }
};

// --- fast form ------------------------------------------------------------
// The same Hamiltonian with per-ray constants folded, one sincos(theta) and
// reciprocals of Delta, Sigma, sin^2 and max(sin, 1e-4) per evaluation, and
// the drift gauges, the step length, the metric factor and the direction of
// a state taken from the evaluation that starts the next step.

struct KerrConsts {
    float a, a2, E, L, aE, aL, L2, a2E2, two_aEL, rplus;
};

GARG_DEV KerrConsts kerr_consts(float a, float E, float L) {
    KerrConsts k;
    k.a = a;
    k.a2 = a * a;
    k.E = E;
    k.L = L;
    k.aE = a * E;
    k.aL = a * L;
    k.L2 = L * L;
    k.a2E2 = k.a2 * E * E;
    k.two_aEL = 2.0f * a * E * L;
    k.rplus = kerr_outer_horizon(a);
    return k;
}

struct KerrEval {
    float dr, dth, dph, dpr, dpth;  // Hamilton's equations at the state
    float st, ct;                   // sin, cos of theta
    float null_residual;            // 2 Sigma H
    float carter;                   // Carter constant
    float inv_sigma;
    float inv_r2;                   // 1 / max(r^2, 1) (GAUGES evaluations only)
};

GARG_DEV float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
GARG_DEV float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

GARG_DEV f3 kerr_position_fast(float a2, float r, float st, float ct, float phi) {
    float sp, cp;
    fast_sincos(phi, &sp, &cp);
    float radial = fast_sqrt(r * r + a2);
    return f3{radial * st * cp, radial * st * sp, r * ct};
}

// The locally measured direction from the sines and cosines of theta and phi.
GARG_DEV f3 kerr_direction_sc(float r, float st, float ct, float sp, float cp, float rdot, float thdot, float phdot) {
    f3 r_hat{st * cp, st * sp, ct};
    f3 theta_hat{ct * cp, ct * sp, -st};
    f3 phi_hat{-sp, cp, 0.0f};
    f3 local = r_hat * rdot + theta_hat * (r * thdot) + phi_hat * (r * st * phdot);
    float len2 = dot(local, local);
    return len2 > 0.0f ? local * __builtin_amdgcn_rsqf(len2) : r_hat;
}

GARG_DEV f3 kerr_direction_fast(float r, float st, float ct, float phi, float rdot, float thdot, float phdot) {
    float sp, cp;
    fast_sincos(phi, &sp, &cp);
    return kerr_direction_sc(r, st, ct, sp, cp, rdot, thdot, phdot);
}

// The right-hand sides at a state whose sin and cos of theta are given;
// GAUGES also yields 1 / max(r^2, 1) for the null-residual gauge.
template <bool GAUGES = false>
GARG_DEV KerrEval kerr_eval_sc(const KerrConsts& k, float r, float st, float ct, float pr, float pth) {
    KerrEval e;
    e.st = st;
    e.ct = ct;
    float r2 = r * r;
    float delta = (r2 - 2.0f * r) + k.a2;
    float sigma = r2 + k.a2 * (e.ct * e.ct);
    // max(sin^2, 1e-8) = max(|sin|, 1e-4)^2: one reciprocal of m serves both
    // clamped terms; max(sin, 1e-4) differs from max(|sin|, 1e-4) only past
    // the pole, where sin < 0. 1/Sigma, 1/Delta and 1/m (and for GAUGES
    // 1/max(r^2, 1)) come from one reciprocal of their product: transcendental
    // instructions issue at a fraction of the VALU rate.
    float m = fmaxf(fabsf(e.st), 1e-4f);
    float sin2 = m * m;
    float sd = sigma * delta;
    float M = GAUGES ? fmaxf(r2, 1.0f) : 1.0f;
    float R = fast_rcp(sd * m * M);
    float inv_m = GAUGES ? sd * M * R : sd * R;
    float mR = GAUGES ? m * M * R : m * R;
    float inv_sigma = delta * mR;
    float inv_delta = sigma * mR;
    e.inv_r2 = GAUGES ? sd * m * R : 1.0f;
    float inv_sin2 = inv_m * inv_m;
    float inv_ms = e.st >= 1e-4f ? inv_m : 1e4f;
    float P = (r2 + k.a2) * k.E - k.aL;
    float P_over_delta = P * inv_delta;
    float G = (k.L2 * inv_sin2 - k.two_aEL) + k.a2E2 * sin2;
    float pr2 = pr * pr, pth2 = pth * pth;
    float F = ((delta * pr2 + pth2) - P * P_over_delta) + G;
    // The factors 1/2, 2 and 4 of H = F / (2 Sigma), dDelta/dr = 2 (r - 1)
    // and the r and theta derivatives of F ride on the half-quantities
    // below: scaling by a power of two is exact in float, so each product
    // equals the one with the factor written out, with fewer instructions.
    float twice_H = F * inv_sigma;
    float half_delta_prime = r - 1.0f;
    float half_fr = (half_delta_prime * pr2 - 2.0f * r * k.E * P_over_delta) +
                    P_over_delta * P_over_delta * half_delta_prime;
    float sc = e.st * e.ct;
    float half_fth = -k.L2 * e.ct * inv_sin2 * inv_ms + k.a2E2 * sc;
    e.dr = delta * pr * inv_sigma;
    e.dth = pth * inv_sigma;
    e.dph = ((k.a * P_over_delta + k.L * inv_sin2) - k.aE) * inv_sigma;
    e.dpr = (r * twice_H - half_fr) * inv_sigma;
    e.dpth = (-k.a2 * sc * twice_H - half_fth) * inv_sigma;
    e.null_residual = F;
    e.carter = pth2 + (e.ct * e.ct) * (k.L2 * inv_sin2 - k.a2E2);
    e.inv_sigma = inv_sigma;
    return e;
}

GARG_DEV KerrEval kerr_eval(const KerrConsts& k, float r, float th, float pr, float pth) {
    float st, ct;
    fast_sincos(th, &st, &ct);
    return kerr_eval_sc<true>(k, r, st, ct, pr, pth);
}

// A Runge-Kutta stage at theta + d from the evaluation `at` of theta: the
// stage's sin and cos come from `at`'s by angle addition (sincos_shift).
GARG_DEV KerrEval kerr_eval_stage(const KerrConsts& k, const KerrEval& at, float r, float th, float d, float pr,
                                  float pth) {
    float st, ct;
    sincos_shift(at.st, at.ct, d, th + d, &st, &ct);
    return kerr_eval_sc(k, r, st, ct, pr, pth);
}

// What the fast Kerr loop reads besides the ray: a by-value block, so a
// caller whose wave shares one request can hold it in scalar registers.
struct KerrLoop {
    uint32_t step_cap;
    float step_factor, step_max, escape_radius;
    float opacity;
    uint32_t redshift;
    uint32_t compensated;
    uint32_t gather;  // sample the volume (non-zero only when there is one)
    // The request's spin and the constants of the Kerr-Schild form that
    // follow from it: a^2, the horizons r+ and r-, a / (r+ - r-). Formed
    // here, so that a wave-uniform loop holds them in scalar registers.
    float spin, a2, rp, rm, shift_scale;
    VolumeView volume;
};

GARG_DEV KerrLoop kerr_loop(const garg_march_uniforms& u, const Scene& sc, bool gather, bool compensated) {
    KerrLoop l;
    l.step_cap = u.geodesic.step_cap;
    l.step_factor = u.geodesic.step_factor;
    l.step_max = u.geodesic.step_max;
    l.escape_radius = u.geodesic.escape_radius;
    l.opacity = u.opacity_scale;
    l.redshift = u.redshift;
    l.compensated = compensated ? 1u : 0u;
    l.gather = gather && sc.volume.voxels != nullptr ? 1u : 0u;
    l.spin = u.spin;
    l.a2 = u.spin * u.spin;
    const float root = sqrtf(fmaxf(1.0f - l.a2, 0.0f));
    l.rp = 1.0f + root;
    l.rm = 1.0f - root;
    l.shift_scale = root > 0.0f ? u.spin / (l.rp - l.rm) : 0.0f;
    l.volume = sc.volume;
    return l;
}

struct KerrFastMarch {
    KerrConsts k;
    float r, th, ph, pr, pth;
    float cr, cth, cph, cpr, cpth;  // TwoSum carries
    KerrEval e;                      // evaluation at the current state
    float carter0, inv_scale, capture;
    Gather g;
    float prev_rho, prev_z, prev_ph;  // the previous point, cylindrical
    float drift;
    uint32_t steps;
    uint32_t outcome;
    bool done;

    GARG_DEV void begin(float spin, const KerrRayState& s) {
        k = kerr_consts(spin, s.energy, s.angular_momentum);
        r = s.r;
        th = s.theta;
        ph = s.phi;
        pr = s.pr;
        pth = s.ptheta;
        cr = cth = cph = cpr = cpth = 0.0f;
        e = kerr_eval(k, r, th, pr, pth);
        carter0 = e.carter;
        inv_scale = 1.0f / fmaxf(fabsf(carter0), k.E * k.E);
        capture = k.rplus + GARG_F(KERR_CAPTURE_MARGIN);
        g.reset();
        prev_rho = fast_sqrt(r * r + k.a2) * e.st;
        prev_z = r * e.ct;
        prev_ph = ph;
        drift = 0.0f;
        steps = 0;
        outcome = GARG_RAY_EXHAUSTED;
        done = false;
    }

    // Continue a ray from state `s` at the point it has reached: the drift
    // reference, the drift so far, the step count and the gather carry over.
    GARG_DEV void resume(float spin, const KerrRayState& s, float carter_ref, float scale_inv, const Gather& gathered,
                         float drift_so_far, uint32_t steps_so_far) {
        k = kerr_consts(spin, s.energy, s.angular_momentum);
        r = s.r;
        th = s.theta;
        ph = s.phi;
        pr = s.pr;
        pth = s.ptheta;
        cr = cth = cph = cpr = cpth = 0.0f;
        e = kerr_eval(k, r, th, pr, pth);
        carter0 = carter_ref;
        inv_scale = scale_inv;
        capture = k.rplus + GARG_F(KERR_CAPTURE_MARGIN);
        g = gathered;
        prev_rho = fast_sqrt(r * r + k.a2) * e.st;
        prev_z = r * e.ct;
        prev_ph = ph;
        drift = drift_so_far;
        steps = steps_so_far;
        outcome = GARG_RAY_EXHAUSTED;
        done = false;
    }

    GARG_DEV void run(const garg_march_uniforms& u, const Scene& sc, bool gather, bool compensated, uint32_t budget) {
        run_loop<false>(kerr_loop(u, sc, gather, compensated), budget);
    }

    // Advance by at most `budget` steps. UNIFORM: every lane of the wave
    // passes the same `l`, which is then read from the first lane only.
    // `tick` runs once per step of the wave (scalar work only: it sees the
    // wave, not the lane).
    template <bool UNIFORM, class Tick = NoTick>
    GARG_DEV void run_loop(KerrLoop l, uint32_t budget, Tick tick = Tick{}) {
        if constexpr (UNIFORM) l = wave_uniform(l);
        for (uint32_t n = 0; n < budget && !done; ++n) {
            tick(n);
            step(l);
        }
        if (!done && steps >= l.step_cap) done = true;
    }

    // One RK4 step with its drift gauges, path and gather; sets `done` at
    // the cap, at capture and at escape.
    GARG_DEV void step(const KerrLoop& l) {
        const bool has_volume = l.gather != 0u;
        const bool compensated = l.compensated != 0u;
        {
            if (steps >= l.step_cap) {
                done = true;
                return;
            }
            float near_h = 0.2f * (r - k.rplus);
            float pole = fminf(fmaxf(e.st * (1.0f / GARG_F(KERR_POLE_SINE)), 0.05f), 1.0f);
            float h = fminf(fmaxf(fminf(l.step_factor * r, near_h) * pole, GARG_F(KERR_STEP_FLOOR)), l.step_max);
            float hh = 0.5f * h;
            KerrEval k2 = kerr_eval_stage(k, e, r + e.dr * hh, th, e.dth * hh, pr + e.dpr * hh, pth + e.dpth * hh);
            KerrEval k3 = kerr_eval_stage(k, e, r + k2.dr * hh, th, k2.dth * hh, pr + k2.dpr * hh, pth + k2.dpth * hh);
            KerrEval k4 = kerr_eval_stage(k, e, r + k3.dr * h, th, k3.dth * h, pr + k3.dpr * h, pth + k3.dpth * h);
            float w = h * (1.0f / 6.0f);
            // Compensated always; without compensation the carries stay zero,
            // which makes each update the plain s + increment.
            fast_two_sum(r, cr, (e.dr + k2.dr * 2.0f) + (k3.dr * 2.0f + k4.dr), w);
            fast_two_sum(th, cth, (e.dth + k2.dth * 2.0f) + (k3.dth * 2.0f + k4.dth), w);
            fast_two_sum(ph, cph, (e.dph + k2.dph * 2.0f) + (k3.dph * 2.0f + k4.dph), w);
            fast_two_sum(pr, cpr, (e.dpr + k2.dpr * 2.0f) + (k3.dpr * 2.0f + k4.dpr), w);
            fast_two_sum(pth, cpth, (e.dpth + k2.dpth * 2.0f) + (k3.dpth * 2.0f + k4.dpth), w);
            if (!compensated) cr = cth = cph = cpr = cpth = 0.0f;
            steps += 1;
            e = kerr_eval(k, r, th, pr, pth);
            // The step's chord in cylindrical coordinates (rho, phi, z):
            // |dp|^2 = drho^2 + 2 rho rho' (1 - cos dphi) + dz^2, where dphi
            // is the exact difference of the two float angles, so the sine
            // and cosine of phi are needed only where the gas is sampled.
            const float rho = fast_sqrt(r * r + k.a2) * e.st;
            const float z = r * e.ct;
            const float drho = rho - prev_rho, dz = z - prev_z;
            const float chord = 2.0f * rho * prev_rho * versine(ph - prev_ph);
            float path = fast_sqrt((drho * drho + chord) + dz * dz);
            g.travelled += path;
            float carter_drift = fabsf(e.carter - carter0) * inv_scale;
            float null_drift = fabsf(e.null_residual) * inv_scale * e.inv_r2;
            drift = fmaxf(drift, fmaxf(carter_drift, null_drift));
            if (r <= capture) {
                outcome = GARG_RAY_CAPTURED;
                done = true;
                return;
            }
            if (has_volume && fabsf(z) < GARG_F(DISK_SLAB_HALF_HEIGHT)) {
                float sp, cp;
                fast_sincos(ph, &sp, &cp);
                const f3 p{rho * cp, rho * sp, z};
                float f = 1.0f - 2.0f * r * e.inv_sigma;
                const KerrEval& ee = e;
                float rr = r;
                g.sample<true>(l.volume, l.opacity, l.redshift != 0u, p, path, fmaxf(f, 0.0f),
                         [&]() { return kerr_direction_sc(rr, ee.st, ee.ct, sp, cp, ee.dr, ee.dth, ee.dph); });
            }
            prev_rho = rho;
            prev_z = z;
            prev_ph = ph;
            if (r >= l.escape_radius) {
                outcome = GARG_RAY_ESCAPED;
                done = true;
                return;
            }
        }
    }

    GARG_DEV bool falls_in() const {
        return outcome == GARG_RAY_CAPTURED || (outcome == GARG_RAY_EXHAUSTED && pr < 0.0f);
    }
    GARG_DEV f3 final_direction() const { return kerr_direction_fast(r, e.st, e.ct, ph, e.dr, e.dth, e.dph); }
    GARG_DEV f3 final_position() const { return kerr_position_fast(k.a2, r, e.st, e.ct, ph); }
};

// ---------------------------------------------------------------------------
// Kerr in Cartesian Kerr-Schild coordinates: the integrator form the
// reference names as future work for the polar seam of its Boyer-Lindquist
// march (docs/lab/07-kerr.md:33-45). The metric g = eta + f l l, with
// f = 2 r^3 / (r^4 + a^2 z^2) and l = (1, (r x + a y) / (r^2 + a^2),
// (r y - a x) / (r^2 + a^2), z / r), is regular on the spin axis, so rays
// that pass over the pole need neither the clamped 1/sin^2 terms nor the
// pole factor of the step length. A photon with p_t = -E moves by
//   dx/dlambda = p - f (E + l.p) l,
//   dp/dlambda = (1/2) (E + l.p)^2 grad f + f (E + l.p) (grad l) p,
// in the affine parameter of the Boyer-Lindquist form. Rays launch from the
// reference's static observer (kerr_launch, converted), and the gas, the
// path lengths and the sky are taken at the Boyer-Lindquist position and
// direction of each point, so the image is the reference's wherever its
// coordinates are regular.

struct KsState {
    float x, y, z, px, py, pz;
};

struct KsConsts {
    float a, a2, E;
    float rp, rm, shift_scale;  // horizons and a / (r+ - r-)
    float phase;                // ln |(r - r+) / (r - r-)| at the launch radius
};

GARG_DEV float ks_radius(float a2, float x, float y, float z) {
    float half = 0.5f * (((x * x + y * y) + z * z) - a2);
    return __builtin_amdgcn_sqrtf(half + __builtin_amdgcn_sqrtf(half * half + a2 * (z * z)));
}

GARG_DEV float ks_log_ratio(const KsConsts& k, float r) { return __logf(fabsf((r - k.rp) / (r - k.rm))); }

// The constants of a ray: the spin's (the same for every ray of a request)
// and the ray's energy and launch phase.
GARG_DEV KsConsts ks_consts(float spin, float E, float phase) {
    KsConsts k;
    k.a = spin;
    k.a2 = spin * spin;
    k.E = E;
    const float root = sqrtf(fmaxf(1.0f - k.a2, 0.0f));
    k.rp = 1.0f + root;
    k.rm = 1.0f - root;
    k.shift_scale = root > 0.0f ? spin / (k.rp - k.rm) : 0.0f;
    k.phase = phase;
    return k;
}

// Hamilton's equations at a state, with the pieces of the metric that the
// drift gauges reuse.
struct KsEval {
    KsState d;
    float r, ir, q, f, big;
};

GARG_DEV KsEval ks_eval(const KsConsts& k, const KsState& s) {
    const float a = k.a, a2 = k.a2;
    KsEval e;
    // r^2 = half + sqrt(half^2 + a^2 z^2); r and 1/r from one reciprocal
    // square root, 1/(r^2 + a^2) and 1/d from one reciprocal.
    const float half = 0.5f * (((s.x * s.x + s.y * s.y) + s.z * s.z) - a2);
    const float rr2 = half + __builtin_amdgcn_sqrtf(half * half + a2 * (s.z * s.z));
    const float ir = __builtin_amdgcn_rsqf(rr2);
    const float r = rr2 * ir;
    const float r2 = r * r;
    const float q = r2 + a2;
    const float d = r2 * r2 + a2 * (s.z * s.z);
    const float iqd = __builtin_amdgcn_rcpf(q * d);
    const float iq = d * iqd, id = q * iqd;
    const float r3 = r2 * r;
    const float f = 2.0f * r3 * id;
    const float lx = (r * s.x + a * s.y) * iq, ly = (r * s.y - a * s.x) * iq, lz = s.z * ir;
    const float big = k.E + ((lx * s.px + ly * s.py) + lz * s.pz);
    const float fb = f * big;
    e.d.x = s.px - fb * lx;
    e.d.y = s.py - fb * ly;
    e.d.z = s.pz - fb * lz;
    // grad r = (x r^3, y r^3, z r (r^2 + a^2)) / d. With f_r = df/dr at
    // fixed z and A = (dl/dr at fixed x, y, z) . p, the momentum equation
    // is (1/2) big^2 (f_r grad r + f_z e_z) + f big (A grad r + (dl/dx)|_r p),
    // so the grad r terms share one factor.
    const float id2 = id * id;
    const float fr = 2.0f * r2 * (3.0f * a2 * (s.z * s.z) - r2 * r2) * id2;
    const float fz = -4.0f * a2 * s.z * r3 * id2;
    const float iq2 = iq * iq;
    const float lrx = (s.x * (a2 - r2) - 2.0f * a * r * s.y) * iq2;
    const float lry = (s.y * (a2 - r2) + 2.0f * a * r * s.x) * iq2;
    const float lrz = -s.z * ir * ir;
    const float A = (lrx * s.px + lry * s.py) + lrz * s.pz;
    const float hb = 0.5f * big * big;
    const float g = (hb * fr + fb * A) * id;
    const float gxy = g * r3, fbq = fb * iq;
    e.d.px = gxy * s.x + fbq * (r * s.px - a * s.py);
    e.d.py = gxy * s.y + fbq * (a * s.px + r * s.py);
    e.d.pz = g * (s.z * r * q) + (hb * fz + fb * s.pz * ir);
    e.r = r;
    e.ir = ir;
    e.q = q;
    e.f = f;
    e.big = big;
    return e;
}

GARG_DEV KsState ks_axpy(const KsState& s, const KsState& k, float h) {
    return KsState{s.x + k.x * h, s.y + k.y * h, s.z + k.z * h, s.px + k.px * h, s.py + k.py * h, s.pz + k.pz * h};
}

// The reference's static-observer launch (Boyer-Lindquist covariant
// momenta) in Kerr-Schild coordinates: raise with the Boyer-Lindquist
// inverse metric, map the vector through the coordinate change (phi~ = phi
// at the launch), lower with the Kerr-Schild metric. Once per ray, in
// precise arithmetic.
GARG_DEV KsState ks_from_bl(const KsConsts& k, const KerrRayState& s) {
    const float a = k.a, a2 = k.a2, r = s.r;
    float st, ct, sp, cp;
    sincosf(s.theta, &st, &ct);
    sincosf(s.phi, &sp, &cp);
    const float sin2 = fmaxf(st * st, 1e-12f);
    const float delta = r * r - 2.0f * r + a2, sigma = r * r + a2 * ct * ct;
    const float E = s.energy, L = s.angular_momentum;
    const float g_tt = -((r * r + a2) * (r * r + a2) - a2 * delta * sin2) / (sigma * delta);
    const float g_tphi = -2.0f * a * r / (sigma * delta);
    const float g_phiphi = (delta - a2 * sin2) / (sigma * delta * sin2);
    const float kt = g_tt * (-E) + g_tphi * L;
    const float kphi = g_tphi * (-E) + g_phiphi * L;
    const float kr = delta * s.pr / sigma, kth = s.ptheta / sigma;
    const float kt_ks = kt + 2.0f * r / delta * kr;
    const float kphi_ks = kphi + a / delta * kr;
    const float x = (r * cp - a * sp) * st, y = (r * sp + a * cp) * st, z = r * ct;
    const float vx = cp * st * kr + (r * cp - a * sp) * ct * kth - y * kphi_ks;
    const float vy = sp * st * kr + (r * sp + a * cp) * ct * kth + x * kphi_ks;
    const float vz = ct * kr - r * st * kth;
    const float rr = ks_radius(a2, x, y, z);
    const float q = rr * rr + a2;
    const float f = 2.0f * rr * rr * rr / (rr * rr * rr * rr + a2 * z * z);
    const float lx = (rr * x + a * y) / q, ly = (rr * y - a * x) / q, lz = z / rr;
    const float lk = kt_ks + (lx * vx + ly * vy) + lz * vz;
    return KsState{x, y, z, vx + f * lx * lk, vy + f * ly * lk, vz + f * lz * lk};
}

// The renderer draws Boyer-Lindquist positions, sqrt(r^2 + a^2)
// sin(theta) (cos phi, sin phi) and r cos(theta): the Kerr-Schild point
// turned about z by -psi(r), psi = atan2(a, r) + (phi~ - phi)(r), whose
// second term is (a / (r+ - r-)) ln |(r - r+) / (r - r-)| less its launch
// value. The march tracks only how psi changes from step to step (a
// small-angle atan and a log1p of the exact differences, no atan2), which
// is what the length of a step's chord needs; the absolute rotation is
// formed only where the gas is sampled and at the end of the ray.
struct KerrSchildMarch {
    // Per ray: the spin (for the finishing calls), the energy and the launch
    // phase; the loop takes the spin's constants from its scalar parameters.
    float a, E, phase;
    KsState s, c;   // state and Fast2Sum carries
    KsEval e;       // evaluation at the current state
    float carter0, l0, inv_scale, inv_root_scale;
    Gather g;
    f3 prev;        // Kerr-Schild position of the previous point
    float prev_r;
    float drift;
    uint32_t steps;
    uint32_t outcome;
    bool done;

    GARG_DEV KsConsts consts() const { return ks_consts(a, E, phase); }

    GARG_DEV void begin(float spin, const KerrRayState& launched) {
        a = spin;
        E = launched.energy;
        KsConsts k = ks_consts(spin, E, 0.0f);
        phase = ks_log_ratio(k, launched.r);
        k.phase = phase;
        s = ks_from_bl(k, launched);
        c = KsState{0, 0, 0, 0, 0, 0};
        e = ks_eval(k, s);
        l0 = s.x * s.py - s.y * s.px;
        carter0 = carter(k);
        const float scale = fmaxf(fabsf(carter0), E * E);
        inv_scale = 1.0f / scale;
        inv_root_scale = 1.0f / sqrtf(scale);
        g.reset();
        prev = f3{s.x, s.y, s.z};
        prev_r = e.r;
        drift = 0.0f;
        steps = 0;
        outcome = GARG_RAY_EXHAUSTED;
        done = false;
    }

    // The Carter constant as the Boyer-Lindquist form writes it,
    // p_theta^2 + cos^2(theta) (L^2 / sin^2(theta) - a^2 E^2), with
    // L = x p_y - y p_x, sin^2(theta) = rho^2 / (r^2 + a^2) and
    // p_theta = (cos(theta) (r^2 + a^2) (x p_x + y p_y) - r rho^2 p_z)
    // / (rho sqrt(r^2 + a^2)). Near the axis, 1 - cos^2(theta) would lose
    // the small sin^2(theta) that L^2 / sin^2(theta) divides by.
    GARG_DEV float carter(const KsConsts& k) const {
        const float rho2 = fmaxf(s.x * s.x + s.y * s.y, 1e-30f);
        const float ct = s.z * e.ir;
        const float rs = __builtin_amdgcn_rsqf(rho2 * e.q);  // 1 / (rho sqrt(r^2 + a^2))
        const float pth = (ct * e.q * (s.x * s.px + s.y * s.py) - e.r * rho2 * s.pz) * rs;
        const float L = s.x * s.py - s.y * s.px;
        // L^2 (r^2 + a^2) / rho^2 = (L (r^2 + a^2) rs)^2
        const float Lq = L * e.q * rs;
        return pth * pth + ct * ct * (Lq * Lq - k.a2 * k.E * k.E);
    }

    // cos and sin of psi at the current point.
    GARG_DEV void rotation(const KsConsts& k, float* cpsi, float* spsi) const {
        const float rq = __builtin_amdgcn_rsqf(e.q);
        const float ca = e.r * rq, sa = k.a * rq;
        float sx, cx;
        fast_sincos(k.shift_scale * (ks_log_ratio(k, e.r) - k.phase), &sx, &cx);
        *cpsi = ca * cx - sa * sx;
        *spsi = sa * cx + ca * sx;
    }

    GARG_DEV f3 position_bl(const KsConsts& k) const {
        float cs, sn;
        rotation(k, &cs, &sn);
        return f3{cs * s.x + sn * s.y, cs * s.y - sn * s.x, s.z};
    }

    // The locally measured direction in the Boyer-Lindquist frame (the
    // convention of kerr_direction) from the coordinate velocities at the
    // current point, whose Boyer-Lindquist position is `p`.
    GARG_DEV f3 direction_bl(const KsConsts& k, f3 p) const {
        const KsState& v = e.d;
        const float r = e.r, r2 = r * r, q = e.q;
        const float d = r2 * r2 + k.a2 * s.z * s.z;
        const float rdot = ((s.x * v.x + s.y * v.y) * r2 * r + s.z * v.z * r * q) / d;
        const float ct = fminf(fmaxf(s.z * e.ir, -1.0f), 1.0f);
        const float st = sqrtf(fmaxf(1.0f - ct * ct, 1e-12f));
        const float rho2 = fmaxf(s.x * s.x + s.y * s.y, 1e-30f);
        const float irho = __builtin_amdgcn_rsqf(rho2);
        // theta' from z = r cos(theta) away from the axis, from
        // rho = sqrt(r^2 + a^2) sin(theta) near it, where the first divides
        // two small numbers.
        float thdot;
        if (fabsf(ct) < 0.7f) {
            thdot = (ct * rdot - v.z) / (r * st);
        } else {
            const float rq = sqrtf(q);
            const float rhodot = (s.x * v.x + s.y * v.y) * irho;
            thdot = (rhodot - r * rdot * st / rq) / (rq * ct);
        }
        const float delta = r2 - 2.0f * r + k.a2;
        const float phdot = (s.x * v.y - s.y * v.x) / rho2 + k.a * rdot / q - k.a * rdot / delta;
        const float cp = p.x * irho, sp = p.y * irho;
        const f3 r_hat{st * cp, st * sp, ct};
        const f3 theta_hat{ct * cp, ct * sp, -st};
        const f3 phi_hat{-sp, cp, 0.0f};
        const f3 local = r_hat * rdot + theta_hat * (r * thdot) + phi_hat * (r * st * phdot);
        const float len2 = dot(local, local);
        return len2 > 0.0f ? local * __builtin_amdgcn_rsqf(len2) : r_hat;
    }

    // Continue a ray from Kerr-Schild state `state` whose azimuth equals
    // the Boyer-Lindquist one at radius `r_phase` of its own phase: the
    // drift references (Carter constant, axial angular momentum, scale), the
    // drift so far, the step count and the gather carry over.
    GARG_DEV void resume(float spin, float energy, const KsState& state, float launch_phase, float carter_ref,
                         float l_ref, float scale_inv, const Gather& gathered, float drift_so_far,
                         uint32_t steps_so_far) {
        a = spin;
        E = energy;
        phase = launch_phase;
        const KsConsts k = consts();
        s = state;
        c = KsState{0, 0, 0, 0, 0, 0};
        e = ks_eval(k, s);
        l0 = l_ref;
        carter0 = carter_ref;
        inv_scale = scale_inv;
        inv_root_scale = sqrtf(scale_inv);
        g = gathered;
        prev = f3{s.x, s.y, s.z};
        prev_r = e.r;
        drift = drift_so_far;
        steps = steps_so_far;
        outcome = GARG_RAY_EXHAUSTED;
        done = false;
    }

    GARG_DEV void run(const garg_march_uniforms& u, const Scene& sc, bool gather, bool compensated, uint32_t budget) {
        run_loop<false>(kerr_loop(u, sc, gather, compensated), budget);
    }

    // The constants of this ray with the spin's taken from the loop's
    // (scalar) parameters.
    GARG_DEV KsConsts loop_consts(const KerrLoop& l) const {
        return KsConsts{l.spin, l.a2, E, l.rp, l.rm, l.shift_scale, phase};
    }

    // Advance by at most `budget` steps (the contract of KerrFastMarch).
    template <bool UNIFORM, class Tick = NoTick>
    GARG_DEV void run_loop(KerrLoop l, uint32_t budget, Tick tick = Tick{}) {
        if constexpr (UNIFORM) l = wave_uniform(l);
        const KsConsts k = loop_consts(l);
        for (uint32_t n = 0; n < budget && !done; ++n) {
            tick(n);
            step(l, k);
        }
        if (!done && steps >= l.step_cap) done = true;
    }

    // One RK4 step with its drift gauges, path and gather; sets `done` at
    // the cap, at capture and at escape.
    GARG_DEV void step(const KerrLoop& l, const KsConsts& k) {
        const bool has_volume = l.gather != 0u;
        const bool compensated = l.compensated != 0u;
        const float capture = k.rp + GARG_F(KERR_CAPTURE_MARGIN);
        {
            if (steps >= l.step_cap) {
                done = true;
                return;
            }
            const float r = e.r;
            // The Boyer-Lindquist step policy without its pole factor.
            const float h = fminf(fmaxf(fminf(l.step_factor * r, 0.2f * (r - k.rp)), GARG_F(KERR_STEP_FLOOR)),
                                  l.step_max);
            const float hh = 0.5f * h;
            // RK4, the increments summed as the stages produce them
            // (e.d + 2 k2 + 2 k3 + k4), so no stage outlives the next.
            KsState sum, kn;
            kn = ks_eval(k, ks_axpy(s, e.d, hh)).d;
            sum = ks_axpy(e.d, kn, 2.0f);
            kn = ks_eval(k, ks_axpy(s, kn, hh)).d;
            sum = ks_axpy(sum, kn, 2.0f);
            kn = ks_eval(k, ks_axpy(s, kn, h)).d;
            const float w = h * (1.0f / 6.0f);
            // Compensated always; without compensation the carries stay
            // zero, which makes each update the plain s + increment.
            fast_two_sum(s.x, c.x, sum.x + kn.x, w);
            fast_two_sum(s.y, c.y, sum.y + kn.y, w);
            fast_two_sum(s.z, c.z, sum.z + kn.z, w);
            fast_two_sum(s.px, c.px, sum.px + kn.px, w);
            fast_two_sum(s.py, c.py, sum.py + kn.py, w);
            fast_two_sum(s.pz, c.pz, sum.pz + kn.pz, w);
            if (!compensated) c = KsState{0, 0, 0, 0, 0, 0};
            steps += 1;
            e = ks_eval(k, s);
            const float r1 = e.r;
            // Drift gauges: the null residual 2 Sigma H everywhere, the Carter
            // constant away from the axis (its formula divides by
            // sin^2(theta)), the axial angular momentum everywhere.
            {
                const float hamilton =
                    0.5f * ((((s.px * s.px + s.py * s.py) + s.pz * s.pz) - k.E * k.E) - e.f * e.big * e.big);
                const float ct = s.z * e.ir;
                const float sigma = r1 * r1 + k.a2 * (ct * ct);
                float dn = fabsf(2.0f * sigma * hamilton) * inv_scale * fminf(e.ir * e.ir, 1.0f);
                if (s.x * s.x + s.y * s.y > 1e-4f * e.q) dn = fmaxf(dn, fabsf(carter(k) - carter0) * inv_scale);
                dn = fmaxf(dn, fabsf((s.x * s.py - s.y * s.px) - l0) * inv_root_scale);
                drift = fmaxf(drift, dn);
            }
            // The chord between the two Boyer-Lindquist points: with
            // delta = psi(r1) - psi(r0),
            // |P1 - P0|^2 = |p1 - p0|^2 + 2 (1 - cos delta) (p0.p1)_xy
            //               - 2 sin(delta) (p0 x p1)_z.
            // atan(a/r1) - atan(a/r0) = atan(a (r0 - r1) / (r0 r1 + a^2)) and
            // the log term's difference is ln(1 + (r1 - r0)(r+ - r-)
            // / ((r1 - r-)(r0 - r+))).
            {
                const float r0 = prev_r, dr = r1 - r0;
                const float u = r0 * r1 + k.a2, v = (r1 - k.rm) * (r0 - k.rp);
                const float iuv = __builtin_amdgcn_rcpf(u * v);
                const float d_atan = small_atan(-k.a * dr * (v * iuv));
                const float d_log = small_log1p(dr * (k.rp - k.rm) * (u * iuv));
                const float delta = d_atan + k.shift_scale * d_log;
                const float dx = s.x - prev.x, dy = s.y - prev.y, dz = s.z - prev.z;
                const float dot_xy = prev.x * s.x + prev.y * s.y;
                const float cross_z = prev.x * s.y - prev.y * s.x;
                const float chord2 = ((dx * dx + dy * dy) + dz * dz) + 2.0f * versine(delta) * dot_xy -
                                     2.0f * small_sin(delta) * cross_z;
                const float path = __builtin_amdgcn_sqrtf(fmaxf(chord2, 0.0f));
                g.travelled += path;
                if (r1 <= capture) {
                    outcome = GARG_RAY_CAPTURED;
                    done = true;
                    return;
                }
                if (has_volume && fabsf(s.z) < GARG_F(DISK_SLAB_HALF_HEIGHT)) {
                    const f3 p = position_bl(k);
                    const float fbl = 1.0f - 2.0f * r1 * __builtin_amdgcn_rcpf(r1 * r1 + k.a2 * (s.z * e.ir) * (s.z * e.ir));
                    g.sample<true>(l.volume, l.opacity, l.redshift != 0u, p, path, fmaxf(fbl, 0.0f),
                             [&]() { return direction_bl(k, p); });
                }
            }
            prev = f3{s.x, s.y, s.z};
            prev_r = r1;
            if (r1 >= l.escape_radius) {
                outcome = GARG_RAY_ESCAPED;
                done = true;
                return;
            }
        }
    }

    // Whether the ray falls into the hole: outward speed negative when the
    // budget ran out, as the Boyer-Lindquist form reads p_r.
    GARG_DEV bool falls_in() const {
        if (outcome == GARG_RAY_CAPTURED) return true;
        if (outcome != GARG_RAY_EXHAUSTED) return false;
        return (s.x * e.d.x + s.y * e.d.y) * e.r * e.r + s.z * e.d.z * e.q < 0.0f;
    }
    GARG_DEV f3 final_direction() const {
        const KsConsts k = consts();
        return direction_bl(k, position_bl(k));
    }
    GARG_DEV f3 final_position() const { return position_bl(consts()); }
};

// ---------------------------------------------------------------------------
// Kerr with Kerr-Schild coordinates only near the spin axis and on fast
// plunges into the horizon. Boyer-Lindquist coordinates are singular on the
// axis, and every ray whose Boyer-Lindquist march breaks the drift budget on
// the 1080p frame at a = 0.9 passes close to it (every such ray dips below
// sin(theta) = 0.2). They are singular on the horizon too: near it the step
// is at most 0.2 (r - r+), so a ray falling faster than dr/dlambda = -5
// (retrograde plungers at a >= 0.93 fall at -5 to -7) takes RK4 substeps
// onto Delta ~ 0, after which its FP32 state is non-finite and it leaves as
// "escaped" one step before the Kerr-Schild march (regular across the
// future horizon) captures it. Rays that graze the horizon and may still
// turn round (prograde, |dr/dlambda| < 1) keep their Boyer-Lindquist steps,
// which hold their drift better than a late switch: handing every ray
// inside r+ + 0.5 to Kerr-Schild raises the frame's largest escaping drift
// from 6.4e-6 to 9.7e-6 and puts 255 rays of the mix's high-spin stills
// over budget. A ray therefore marches in the fast Boyer-Lindquist form
// until sin(theta) first falls below GARG_POLAR_ENTER, or it is inside
// r+ + GARG_HORIZON_ENTER falling faster than GARG_PLUNGE_RATE, and in
// Kerr-Schild coordinates from there to its end. On the a = 0.9 frame 1.1%
// of the steps are below that sin(theta). Both forms take the same step
// length (the pole factor is 1 above sin(theta) = 0.08; the horizon factor
// depends on r alone); the launch's Carter constant stays the drift
// reference and the gather carries across; the coordinate change runs in
// double precision, once per ray that switches.
#define GARG_POLAR_ENTER 0.25f
#define GARG_HORIZON_ENTER 0.5f
#define GARG_PLUNGE_RATE 2.0f

// Boyer-Lindquist (r, theta, phi, p_r, p_theta; E, L) to Kerr-Schild
// (x, y, z, p_x, p_y, p_z), with the Kerr-Schild azimuth equal to phi here:
// ks_from_bl in double precision.
GARG_DEV KsState ks_from_bl_f64(double a, const KerrRayState& st) {
    const double a2 = a * a, r = st.r;
    double sth, cth, sph, cph;
    sincos((double)st.theta, &sth, &cth);
    sincos((double)st.phi, &sph, &cph);
    const double sin2 = fmax(sth * sth, 1e-24);
    const double delta = r * r - 2.0 * r + a2, sigma = r * r + a2 * cth * cth;
    const double E = st.energy, L = st.angular_momentum;
    const double g_tt = -((r * r + a2) * (r * r + a2) - a2 * delta * sin2) / (sigma * delta);
    const double g_tphi = -2.0 * a * r / (sigma * delta);
    const double g_phiphi = (delta - a2 * sin2) / (sigma * delta * sin2);
    const double kt = g_tt * (-E) + g_tphi * L;
    const double kphi = g_tphi * (-E) + g_phiphi * L;
    const double kr = delta * (double)st.pr / sigma, kth = (double)st.ptheta / sigma;
    const double kt_ks = kt + 2.0 * r / delta * kr;
    const double kphi_ks = kphi + a / delta * kr;
    const double x = (r * cph - a * sph) * sth, y = (r * sph + a * cph) * sth, z = r * cth;
    const double vx = cph * sth * kr + (r * cph - a * sph) * cth * kth - y * kphi_ks;
    const double vy = sph * sth * kr + (r * sph + a * cph) * cth * kth + x * kphi_ks;
    const double vz = cth * kr - r * sth * kth;
    const double q = r * r + a2;
    const double f = 2.0 * r * r * r / (r * r * r * r + a2 * z * z);
    const double lx = (r * x + a * y) / q, ly = (r * y - a * x) / q, lz = z / r;
    const double lk = kt_ks + (lx * vx + ly * vy) + lz * vz;
    return KsState{(float)x, (float)y, (float)z, (float)(vx + f * lx * lk), (float)(vy + f * ly * lk),
                   (float)(vz + f * lz * lk)};
}

// A ray's Kerr-Schild state where it leaves the Boyer-Lindquist form, and
// the log of |r - r+| / |r - r-| there, in double precision. Not inlined:
// a ray passes here at most once, and kept apart its double-precision
// constants and temporaries do not stay live in the march's registers.
struct PolarEntry {
    KsState state;
    float log_ratio;
};

__device__ __attribute__((noinline)) PolarEntry polar_entry_f64(float a, KerrRayState st) {
    const double root = sqrt(fmax(1.0 - (double)a * a, 0.0));
    const double ratio = fabs(((double)st.r - (1.0 + root)) / ((double)st.r - (1.0 - root)));
    return PolarEntry{ks_from_bl_f64(a, st), (float)log(ratio)};
}

// The ray is kept in a KerrSchildMarch. While it is in the Boyer-Lindquist
// form that struct's float fields hold the KerrFastMarch fields instead
// (unpack/pack below, plain value copies the compiler keeps in the same
// registers): the shared fields (gather, drift, step count, outcome, done,
// Carter reference, scale) in their own places, E in E, the launch L in l0,
// the state, carries and evaluation in the Kerr-Schild state, carries and
// evaluation, the previous point in prev, and the products of E and L in
// the remaining floats. The spin's constants come from the loop's scalar
// parameters.
struct KerrHybridMarch {
    KerrSchildMarch ks;
    bool polar;  // the Kerr-Schild form carries the ray

    GARG_DEV KerrFastMarch unpack(float a, float a2, float rplus) const {
        KerrFastMarch m;
        m.k.a = a;
        m.k.a2 = a2;
        m.k.E = ks.E;
        m.k.L = ks.l0;
        m.k.aE = ks.phase;
        m.k.aL = ks.inv_root_scale;
        m.k.L2 = ks.prev_r;
        m.k.a2E2 = ks.s.pz;
        m.k.two_aEL = ks.c.pz;
        m.k.rplus = rplus;
        m.r = ks.s.x;
        m.th = ks.s.y;
        m.ph = ks.s.z;
        m.pr = ks.s.px;
        m.pth = ks.s.py;
        m.cr = ks.c.x;
        m.cth = ks.c.y;
        m.cph = ks.c.z;
        m.cpr = ks.c.px;
        m.cpth = ks.c.py;
        m.e.dr = ks.e.d.x;
        m.e.dth = ks.e.d.y;
        m.e.dph = ks.e.d.z;
        m.e.dpr = ks.e.d.px;
        m.e.dpth = ks.e.d.py;
        m.e.st = ks.e.d.pz;
        m.e.ct = ks.e.r;
        m.e.null_residual = ks.e.ir;
        m.e.carter = ks.e.q;
        m.e.inv_sigma = ks.e.f;
        m.e.inv_r2 = ks.e.big;
        m.carter0 = ks.carter0;
        m.inv_scale = ks.inv_scale;
        m.capture = rplus + GARG_F(KERR_CAPTURE_MARGIN);
        m.g = ks.g;
        m.prev_rho = ks.prev.x;
        m.prev_z = ks.prev.y;
        m.prev_ph = ks.prev.z;
        m.drift = ks.drift;
        m.steps = ks.steps;
        m.outcome = ks.outcome;
        m.done = ks.done;
        return m;
    }

    GARG_DEV void pack(const KerrFastMarch& m) {
        ks.a = m.k.a;
        ks.E = m.k.E;
        ks.l0 = m.k.L;
        ks.phase = m.k.aE;
        ks.inv_root_scale = m.k.aL;
        ks.prev_r = m.k.L2;
        ks.s = KsState{m.r, m.th, m.ph, m.pr, m.pth, m.k.a2E2};
        ks.c = KsState{m.cr, m.cth, m.cph, m.cpr, m.cpth, m.k.two_aEL};
        ks.e.d = KsState{m.e.dr, m.e.dth, m.e.dph, m.e.dpr, m.e.dpth, m.e.st};
        ks.e.r = m.e.ct;
        ks.e.ir = m.e.null_residual;
        ks.e.q = m.e.carter;
        ks.e.f = m.e.inv_sigma;
        ks.e.big = m.e.inv_r2;
        ks.carter0 = m.carter0;
        ks.inv_scale = m.inv_scale;
        ks.g = m.g;
        ks.prev = f3{m.prev_rho, m.prev_z, m.prev_ph};
        ks.drift = m.drift;
        ks.steps = m.steps;
        ks.outcome = m.outcome;
        ks.done = m.done;
    }

    GARG_DEV KerrFastMarch unpack() const {
        const float a = ks.a, a2 = a * a;
        return unpack(a, a2, kerr_outer_horizon(a));
    }

    // From the Boyer-Lindquist form to the Kerr-Schild form at the current point.
    GARG_DEV void to_polar(const KerrFastMarch& m) {
        const KerrRayState st{m.r, m.th, m.ph, m.pr, m.pth, m.k.E, m.k.L};
        const PolarEntry entry = polar_entry_f64(m.k.a, st);
        ks.resume(m.k.a, st.energy, entry.state, entry.log_ratio, m.carter0, m.k.L, m.inv_scale, m.g, m.drift, m.steps);
        polar = true;
    }

    GARG_DEV void begin(float spin, const KerrRayState& launched) {
        KerrFastMarch m;
        m.begin(spin, launched);
        polar = false;
        pack(m);
        if (m.e.st < GARG_POLAR_ENTER || (m.r < m.k.rplus + GARG_HORIZON_ENTER && m.e.dr < -GARG_PLUNGE_RATE)) to_polar(m);
    }

    GARG_DEV bool done() const { return ks.done; }

    GARG_DEV void run(const garg_march_uniforms& u, const Scene& sc, bool gather, bool compensated, uint32_t budget) {
        run_loop<false>(kerr_loop(u, sc, gather, compensated), budget);
    }

    // Two phases, so that only one form's state is live at a time: the
    // Boyer-Lindquist steps until the ray ends, nears the axis or plunges
    // into the horizon, then the Kerr-Schild steps for the rest of the budget.
    template <bool UNIFORM, class Tick = NoTick>
    GARG_DEV void run_loop(KerrLoop l, uint32_t budget, Tick tick = Tick{}) {
        if constexpr (UNIFORM) l = wave_uniform(l);
        uint32_t n = 0;
        if (!polar) {
            KerrFastMarch m = unpack(l.spin, l.a2, l.rp);
            bool axis = false;
            for (; n < budget && !m.done; ++n) {
                tick(n);
                m.step(l);
                if (!m.done && (m.e.st < GARG_POLAR_ENTER || (m.r < l.rp + GARG_HORIZON_ENTER && m.e.dr < -GARG_PLUNGE_RATE))) {
                    axis = true;
                    ++n;
                    break;
                }
            }
            if (axis) to_polar(m);
            else pack(m);
        }
        if (polar) {
            const KsConsts k = ks.loop_consts(l);
            for (; n < budget && !ks.done; ++n) {
                tick(n);
                ks.step(l, k);
            }
        }
        if (!ks.done && ks.steps >= l.step_cap) ks.done = true;
    }

    GARG_DEV const Gather& gather() const { return ks.g; }
    GARG_DEV float max_drift() const { return ks.drift; }
    GARG_DEV uint32_t step_count() const { return ks.steps; }
    GARG_DEV uint32_t result() const { return ks.outcome; }
    GARG_DEV bool falls_in() const { return polar ? ks.falls_in() : unpack().falls_in(); }
    GARG_DEV f3 final_direction() const { return polar ? ks.final_direction() : unpack().final_direction(); }
    GARG_DEV f3 final_position() const { return polar ? ks.final_position() : unpack().final_position(); }
};

// Colour of a finished ray: gas plus the starfield behind it unless it falls
// in (march.metal:226-229, 261-263).
GARG_DEV f3 finish_color(const Gather& g, bool falls_in, f3 sky, const garg_march_uniforms& u, const Scene& sc) {
    f3 color = g.color;
    if (!falls_in && u.star_brightness > 0.0f)
        color = color + starfield(sky, u.star_seed, sc.blackbody) * (g.transmittance * u.star_brightness);
    return color;
}

}  // namespace garg
