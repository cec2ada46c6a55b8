// Particle pass for gfx950 (sim.metal): seeding and semi-implicit Euler in
// the Paczynski-Wiita potential with drag, swept-segment capture and
// annulus respawn. One thread per particle over a structure-of-arrays store,
// so every field is a coalesced 4-byte stream (32 bytes read and 32 written
// per particle per dispatch). Each thread touches only its own particle and
// the library is built with contraction off, so a seed reproduces the store
// bit for bit and a CPU replay of the same float operations matches it.
#include "device_common.hpp"
#include "launch.hpp"

namespace garg {

// d Phi / dr and the circular speed (Common.h:32-46).
GARG_DEV float potential_gradient(float r, uint32_t potential) {
    if (potential == GARG_POTENTIAL_PACZYNSKI_WIITA) {
        float d = fmaxf(r - GARG_F(R_SCHWARZSCHILD), 1e-6f);
        return 1.0f / (d * d);
    }
    return 1.0f / (r * r);
}

// Unnormalized thin disk profile (Common.h:49-53): x^(3/4) t^(1/4) with
// x = r_isco / r and t the clamped torque factor, evaluated as the fourth
// root of x^3 t by two correctly rounded square roots in place of two pow
// calls, which took two fifths of a sim dispatch. Over the particles' radii
// [2.02, 60] it is within 1.2e-7 of the exact value, against 1.4e-7 for
// the product of two rounded pows, and within 2.4e-7 of that product.
GARG_DEV float disk_shape(float r) {
    float x = GARG_F(R_ISCO) / r;
    float torque = fmaxf(1.0f - sqrtf(x), GARG_F(ZERO_TORQUE_FLOOR));
    float q = x * x;
    q = q * x;
    q = q * torque;
    return sqrtf(sqrtf(q));
}

GARG_DEV float len3(float x, float y, float z) {
    float s = x * x;
    s = s + y * y;
    s = s + z * z;
    return sqrtf(s);
}

struct Spawned {
    float px, py, pz, vx, vy, vz, t;
};

// Event 0 of the disk generator draws r from the steady profile's inverse
// CDF; every other event respawns in the feeding annulus (sim.metal:6-34).
GARG_DEV Spawned spawn(const SimArgs& a, uint32_t index, uint32_t event) {
    uint4 h0 = pcg4d(make_uint4(index, event, a.seed_lo, a.seed_hi));
    uint4 h1 = pcg4d(make_uint4(index, event, a.seed_lo ^ 0x9E3779B9u, a.seed_hi));
    float u0 = unit_float(h0.x), u1 = unit_float(h0.y), u2 = unit_float(h0.z), u3 = unit_float(h0.w);
    float u4 = unit_float(h1.x), u5 = unit_float(h1.y);
    float r;
    bool from_disk = event == 0u && a.generator == GARG_GENERATOR_DISK;
    if (from_disk) {
        float position = u0 * (float)(a.inverse_cdf_count - 1u);
        uint32_t k = min((uint32_t)position, a.inverse_cdf_count - 2u);
        float lo = a.inverse_cdf[k], hi = a.inverse_cdf[k + 1u];
        r = lo + (hi - lo) * (position - (float)k);
    } else {
        const float inner = GARG_F(R_FEED_INNER) * GARG_F(R_FEED_INNER);
        const float outer = GARG_F(R_DISK_OUTER) * GARG_F(R_DISK_OUTER);
        r = sqrtf(inner + (outer - inner) * u0);
    }
    const float two_pi = 6.283185307179586f;
    float phi = two_pi * u1;
    float rad01 = sqrtf(-2.0f * logf(1.0f - u2)), ang01 = two_pi * u3;
    float rad23 = sqrtf(-2.0f * logf(1.0f - u4)), ang23 = two_pi * u5;
    float g0 = rad01 * cosf(ang01), g1 = rad01 * sinf(ang01);
    float g2 = rad23 * cosf(ang23), g3 = rad23 * sinf(ang23);
    float vc = sqrtf(r * potential_gradient(r, a.potential));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    float height = (event == 0u && a.generator == GARG_GENERATOR_ANNULUS) ? a.sigma_z * g0
                                                                          : GARG_F(DISK_ASPECT) * r * g0;
    const float disp = GARG_F(FEED_VELOCITY_DISPERSION) * vc;
    Spawned s;
    s.px = r * cp;
    s.py = r * sp;
    s.pz = height;
    s.vx = vc * -sp + disp * g1;
    s.vy = vc * cp + disp * g2;
    s.vz = 0.0f * vc + disp * g3;
    s.t = a.temperature_scale * disk_shape(r);
    return s;
}

__global__ __launch_bounds__(256) void seed_kernel(SimArgs a) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i == 0u) a.p.dispatches[0] = 0u;
    if (i >= a.p.count) return;
    Spawned s = spawn(a, i, 0u);
    a.p.px[i] = s.px;
    a.p.py[i] = s.py;
    a.p.pz[i] = s.pz;
    a.p.vx[i] = s.vx;
    a.p.vy[i] = s.vy;
    a.p.vz[i] = s.vz;
    a.p.temperature[i] = s.t;
    a.p.respawns[i] = 0u;
}

// `dispatches` dispatches of two substeps each of kick, drag, drift;
// capture on the swept chord, escape and the plunge speed limit respawn
// (sim.metal:44-79). Particles do not interact and a dispatch carries only
// position, velocity and the respawn count to the next (the temperature is
// a function of the final position), so consecutive dispatches run in one
// launch with the particle in registers: the same float operations in the
// same order as one launch per dispatch, with one read and one write of
// the store instead of one per dispatch.
__global__ __launch_bounds__(256) void step_kernel(SimArgs a, uint32_t dispatches) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    // One thread counts the dispatches; nothing in this launch reads the count.
    if (i == 0u) a.p.dispatches[0] += dispatches;
    if (i >= a.p.count) return;
    float x = a.p.px[i], y = a.p.py[i], z = a.p.pz[i];
    float vx = a.p.vx[i], vy = a.p.vy[i], vz = a.p.vz[i];
    uint32_t respawns = a.p.respawns[i];
    const float dt = a.dt;
    for (uint32_t s = 0; s < SIM_SUBSTEPS * dispatches; ++s) {
        float r = len3(x, y, z);
        float grad = potential_gradient(r, a.potential);
        float k = -grad / r;
        float omega = sqrtf(r * grad) / r;
        vx = vx + (x * k) * dt;
        vy = vy + (y * k) * dt;
        vz = vz + (z * k) * dt;
        float drag = a.alpha * omega;
        drag = drag * dt;
        vx = vx - vx * drag;
        vy = vy - vy * drag;
        vz = vz - vz * drag;
        float bx = x, by = y, bz = z;
        x = x + vx * dt;
        y = y + vy * dt;
        z = z + vz * dt;
        r = len3(x, y, z);
        float tx = x - bx, ty = y - by, tz = z - bz;
        float tt = tx * tx;
        tt = tt + ty * ty;
        tt = tt + tz * tz;
        float bt = bx * tx;
        bt = bt + by * ty;
        bt = bt + bz * tz;
        float along = -bt / fmaxf(tt, 1e-12f);
        along = fminf(fmaxf(along, 0.0f), 1.0f);
        float closest = len3(bx + along * tx, by + along * ty, bz + along * tz);
        float speed = len3(vx, vy, vz);
        if (closest < GARG_F(R_CAPTURE) || r > GARG_F(R_ESCAPE) || speed > GARG_F(PLUNGE_SPEED_LIMIT)) {
            respawns += 1u;
            Spawned sp = spawn(a, i, respawns);
            x = sp.px;
            y = sp.py;
            z = sp.pz;
            vx = sp.vx;
            vy = sp.vy;
            vz = sp.vz;
        }
    }
    a.p.px[i] = x;
    a.p.py[i] = y;
    a.p.pz[i] = z;
    a.p.vx[i] = vx;
    a.p.vy[i] = vy;
    a.p.vz[i] = vz;
    a.p.temperature[i] = a.temperature_scale * disk_shape(len3(x, y, z));
    a.p.respawns[i] = respawns;
}

__global__ __launch_bounds__(256) void pack_kernel(ParticleView p, garg_particle* out) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.count) return;
    garg_particle r;
    r.position[0] = p.px[i];
    r.position[1] = p.py[i];
    r.position[2] = p.pz[i];
    r.velocity[0] = p.vx[i];
    r.velocity[1] = p.vy[i];
    r.velocity[2] = p.vz[i];
    r.temperature = p.temperature[i];
    r.respawns = p.respawns[i];
    out[i] = r;
}

__global__ __launch_bounds__(256) void unpack_kernel(ParticleView p, const garg_particle* in) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= p.count) return;
    garg_particle r = in[i];
    p.px[i] = r.position[0];
    p.py[i] = r.position[1];
    p.pz[i] = r.position[2];
    p.vx[i] = r.velocity[0];
    p.vy[i] = r.velocity[1];
    p.vz[i] = r.velocity[2];
    p.temperature[i] = r.temperature;
    p.respawns[i] = r.respawns;
}

static dim3 grid_for(uint32_t n) { return dim3((n + 255u) / 256u); }

hipError_t launch_seed(const SimArgs& a, hipStream_t stream) {
    seed_kernel<<<grid_for(a.p.count), 256, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_step(const SimArgs& a, uint32_t dispatches, hipStream_t stream) {
    step_kernel<<<grid_for(a.p.count), 256, 0, stream>>>(a, dispatches);
    return hipGetLastError();
}

hipError_t launch_pack_aos(const ParticleView& p, garg_particle* out, hipStream_t stream) {
    pack_kernel<<<grid_for(p.count), 256, 0, stream>>>(p, out);
    return hipGetLastError();
}

hipError_t launch_unpack_aos(const ParticleView& p, const garg_particle* in, hipStream_t stream) {
    unpack_kernel<<<grid_for(p.count), 256, 0, stream>>>(p, in);
    return hipGetLastError();
}

}  // namespace garg
