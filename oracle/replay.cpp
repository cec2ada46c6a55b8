// Exact single precision replays of the integer deposit and of one sim
// dispatch, for bit-exact parity of the GPU kernels. These restate the
// reference kernels' arithmetic (splat.metal:16-88, sim.metal:44-79) with an
// explicit operation order, compiled with contraction off, so that every
// float operation rounds the same way the GPU kernels round it.
// Test infrastructure only (see oracle.h).
#include "oracle.h"
#include "oracle_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

struct ParticleRecord {
    float position[3];
    float velocity[3];
    float temperature;
    uint32_t respawns;
};
static_assert(sizeof(ParticleRecord) == 32, "particle record is 32 bytes");

inline uint32_t to_u32(float x) { return (uint32_t)(int64_t)x; }

inline void deposit(uint32_t* emission, int32_t* momentum, uint32_t* mass, uint32_t index, float weight, float energy,
                    float relative, const float velocity[3]) {
    const float ef = (float)ENERGY_FIXED, etf = (float)ENERGY_T_FIXED, mf = (float)MOMENTUM_FIXED,
                af = (float)MASS_FIXED;
    float e = energy * weight;
    e = e * ef;
    e = e + 0.5f;
    float et = energy * relative;
    et = et * weight;
    et = et * etf;
    et = et + 0.5f;
    float wm = weight * mf;
    float m = weight * af;
    m = m + 0.5f;
    emission[2u * index + 0u] += to_u32(e);
    emission[2u * index + 1u] += to_u32(et);
    for (int c = 0; c < 3; ++c) {
        float v = velocity[c] * wm;
        momentum[3u * index + (uint32_t)c] += (int32_t)std::rint(v);
    }
    mass[index] += to_u32(m);
}

inline float unit_float(uint32_t x) { return (float)(x >> 8u) * (1.0f / 16777216.0f); }

inline float length3(float x, float y, float z) {
    float s = x * x;
    s = s + y * y;
    s = s + z * z;
    return std::sqrt(s);
}

}  // namespace

extern "C" {

void orc_splat_replay(const void* particles, uint32_t count, uint32_t size, float half_extent, float half_extent_z,
                      float peak_temperature, float emission_scale, uint32_t exact, uint32_t frame, uint64_t seed,
                      uint32_t* emission, int32_t* momentum, uint32_t* mass) {
    const ParticleRecord* ps = (const ParticleRecord*)particles;
    const float half[3] = {half_extent, half_extent, half_extent_z};
    const float span[3] = {2.0f * half_extent, 2.0f * half_extent, 2.0f * half_extent_z};
    const float n = (float)size;
    for (uint32_t i = 0; i < count; ++i) {
        const ParticleRecord& p = ps[i];
        float cube[3];
        bool outside = false;
        for (int a = 0; a < 3; ++a) {
            cube[a] = (p.position[a] + half[a]) / span[a];
            if (cube[a] < 0.0f || cube[a] >= 1.0f) outside = true;
        }
        if (outside) continue;
        float relative = p.temperature / peak_temperature;
        float energy = relative * relative;
        energy = energy * relative;
        energy = energy * relative;
        energy = energy * emission_scale;
        int low[3], high[3];
        float fraction[3];
        for (int a = 0; a < 3; ++a) {
            float grid = cube[a] * n;
            grid = grid - 0.5f;
            float base = std::floor(grid);
            fraction[a] = grid - base;
            low[a] = (int)base;
            high[a] = std::min(low[a] + 1, (int)size - 1);
            low[a] = std::max(low[a], 0);
        }
        if (exact) {
            for (uint32_t c = 0; c < CIC_CORNERS; ++c) {
                int cx = (c & 1u) ? high[0] : low[0];
                int cy = (c & 2u) ? high[1] : low[1];
                int cz = (c & 4u) ? high[2] : low[2];
                float wx = (c & 1u) ? fraction[0] : 1.0f - fraction[0];
                float wy = (c & 2u) ? fraction[1] : 1.0f - fraction[1];
                float wz = (c & 4u) ? fraction[2] : 1.0f - fraction[2];
                float w = wx * wy;
                w = w * wz;
                if (w <= 0.0f) continue;
                uint32_t index = ((uint32_t)cz * size + (uint32_t)cy) * size + (uint32_t)cx;
                deposit(emission, momentum, mass, index, w, energy, relative, p.velocity);
            }
        } else {
            uint32_t in[4] = {i, frame, (uint32_t)seed, (uint32_t)(seed >> 32)}, h[4];
            orc_pcg4d(in, h);
            int cx = unit_float(h[0]) < fraction[0] ? high[0] : low[0];
            int cy = unit_float(h[1]) < fraction[1] ? high[1] : low[1];
            int cz = unit_float(h[2]) < fraction[2] ? high[2] : low[2];
            uint32_t index = ((uint32_t)cz * size + (uint32_t)cy) * size + (uint32_t)cx;
            deposit(emission, momentum, mass, index, 1.0f, energy, relative, p.velocity);
        }
    }
}

void orc_sim_replay_step(void* particles, uint32_t count, int potential, float dt, float alpha, float temperature_scale,
                         uint8_t* respawned) {
    ParticleRecord* ps = (ParticleRecord*)particles;
    const float capture = (float)R_CAPTURE, escape = (float)R_ESCAPE, limit = (float)PLUNGE_SPEED_LIMIT;
    for (uint32_t i = 0; i < count; ++i) {
        ParticleRecord& p = ps[i];
        float x = p.position[0], y = p.position[1], z = p.position[2];
        float vx = p.velocity[0], vy = p.velocity[1], vz = p.velocity[2];
        respawned[i] = 0;
        for (uint32_t s = 0; s < SIM_SUBSTEPS; ++s) {
            float r = length3(x, y, z);
            float grad;
            if (potential == 1) {
                float d = std::fmax(r - (float)R_SCHWARZSCHILD, 1e-6f);
                grad = 1.0f / (d * d);
            } else {
                grad = 1.0f / (r * r);
            }
            float k = -grad / r;
            float omega = std::sqrt(r * grad) / r;
            vx = vx + (x * k) * dt;
            vy = vy + (y * k) * dt;
            vz = vz + (z * k) * dt;
            float drag = alpha * omega;
            drag = drag * dt;
            vx = vx - vx * drag;
            vy = vy - vy * drag;
            vz = vz - vz * drag;
            float bx = x, by = y, bz = z;
            x = x + vx * dt;
            y = y + vy * dt;
            z = z + vz * dt;
            r = length3(x, y, z);
            float tx = x - bx, ty = y - by, tz = z - bz;
            float tt = tx * tx;
            tt = tt + ty * ty;
            tt = tt + tz * tz;
            float bt = bx * tx;
            bt = bt + by * ty;
            bt = bt + bz * tz;
            float along = -bt / std::fmax(tt, 1e-12f);
            along = std::fmin(std::fmax(along, 0.0f), 1.0f);
            float closest = length3(bx + along * tx, by + along * ty, bz + along * tz);
            float speed = length3(vx, vy, vz);
            if (closest < capture || r > escape || speed > limit) {
                respawned[i] = 1;
                break;
            }
        }
        if (respawned[i]) continue;
        p.position[0] = x;
        p.position[1] = y;
        p.position[2] = z;
        p.velocity[0] = vx;
        p.velocity[1] = vy;
        p.velocity[2] = vz;
        float rr = length3(x, y, z);
        float xx = (float)R_ISCO / rr;
        float torque = std::fmax(1.0f - std::sqrt(xx), (float)ZERO_TORQUE_FLOOR);
        p.temperature = temperature_scale * (std::pow(xx, 0.75f) * std::pow(torque, 0.25f));
    }
}

}  // extern "C"
