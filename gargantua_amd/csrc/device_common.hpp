// Device helpers shared by every gfx950 kernel of the library: vector
// arithmetic with a fixed operation order, the pcg4d hash, half storage,
// trilinear volume sampling, table lookups, the starfield, the redshift
// factor and the front-to-back gather.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gargantua.h"
#include "../../include/gargantua_constants.h"

#define GARG_DEV __device__ __forceinline__

namespace garg {

struct f3 {
    float x, y, z;
};

GARG_DEV f3 mk(float x, float y, float z) { return f3{x, y, z}; }
GARG_DEV f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
GARG_DEV f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
GARG_DEV f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
GARG_DEV f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
GARG_DEV f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
// Left to right, as a dot product is written.
GARG_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GARG_DEV float length(f3 a) { return sqrtf(dot(a, a)); }
GARG_DEV f3 normalize(f3 a) { return a / length(a); }
GARG_DEV f3 load3(const float* p) { return f3{p[0], p[1], p[2]}; }

// pcg4d, Jarzynski and Olano 2020, JCGT 9(3), 20 (Common.h:11-17).
GARG_DEV uint4 pcg4d(uint4 v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    v.x ^= v.x >> 16u;
    v.y ^= v.y >> 16u;
    v.z ^= v.z >> 16u;
    v.w ^= v.w >> 16u;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    return v;
}

// Uniform in [0, 1) from the top 24 bits, exact in single precision.
GARG_DEV float unit_float(uint32_t x) { return (float)(x >> 8u) * (1.0f / 16777216.0f); }

typedef _Float16 h4 __attribute__((ext_vector_type(4)));

// One resolved voxel: emission rgb + density, velocity xyz + 1, as halves.
struct Voxel {
    h4 emission;
    h4 velocity;
};
static_assert(sizeof(Voxel) == 16, "voxel is 16 bytes");

struct VolumeView {
    const Voxel* voxels;  // size^3, z-major, x fastest
    int size;
    float half_xy;
    float half_z;
    float scale_xy;  // size / (2 half_xy): texels per unit length
    float scale_z;   // size / (2 half_z)
};

struct Sample {
    f3 emission;
    float density;
    f3 velocity;
};

// One corner's share of a trilinear sample: each of the seven channels
// is added as fma(texel, w, sum), one rounding per channel; a half texel
// and a float sum make one mixed-precision FMA, with no conversion.
GARG_DEV void add_corner(Sample& s, const Voxel& t, float w) {
    s.emission.x = fmaf((float)t.emission.x, w, s.emission.x);
    s.emission.y = fmaf((float)t.emission.y, w, s.emission.y);
    s.emission.z = fmaf((float)t.emission.z, w, s.emission.z);
    s.density = fmaf((float)t.emission.w, w, s.density);
    s.velocity.x = fmaf((float)t.velocity.x, w, s.velocity.x);
    s.velocity.y = fmaf((float)t.velocity.y, w, s.velocity.y);
    s.velocity.z = fmaf((float)t.velocity.z, w, s.velocity.z);
}

// Trilinear filtering on texel centres with zero outside the cube, the
// semantics of the reference's clamp_to_zero sampler (march.metal:15-27) as
// the oracle states them (VolumeField.swift:40-63).
GARG_DEV Sample sample_volume(const VolumeView& v, f3 p) {
    float tx = (p.x + v.half_xy) * v.scale_xy - 0.5f;
    float ty = (p.y + v.half_xy) * v.scale_xy - 0.5f;
    float tz = (p.z + v.half_z) * v.scale_z - 0.5f;
    float bx = floorf(tx), by = floorf(ty), bz = floorf(tz);
    float fx = tx - bx, fy = ty - by, fz = tz - bz;
    int x0 = (int)bx, y0 = (int)by, z0 = (int)bz;
    // The voxels live in global memory: saying so turns generic (flat) loads,
    // which also wait on the LDS counter, into global ones.
    const __attribute__((address_space(1))) Voxel* g = (const __attribute__((address_space(1))) Voxel*)v.voxels;
    Sample s{f3{0, 0, 0}, 0.0f, f3{0, 0, 0}};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int x = x0 + (c & 1), y = y0 + ((c >> 1) & 1), z = z0 + ((c >> 2) & 1);
        float w = ((c & 1) ? fx : 1.0f - fx) * (((c >> 1) & 1) ? fy : 1.0f - fy) * (((c >> 2) & 1) ? fz : 1.0f - fz);
        bool inside = (unsigned)x < (unsigned)v.size && (unsigned)y < (unsigned)v.size && (unsigned)z < (unsigned)v.size;
        if (inside && w != 0.0f) {
            uint32_t i = ((uint32_t)z * (uint32_t)v.size + (uint32_t)y) * (uint32_t)v.size + (uint32_t)x;
            const Voxel t{g[i].emission, g[i].velocity};
            add_corner(s, t, w);
        }
    }
    return s;
}

// sample_volume in two halves, so a caller can keep one sample's texel
// loads in flight while it composites the previous one: fetch_corners
// issues the eight corner loads (from in-range indices; a corner outside
// the volume or of zero weight is dropped by combine_corners, as
// sample_volume skips it), combine_corners forms the same sums in the same
// order as sample_volume.
struct TexelCorners {
    Voxel v[8];
    float fx, fy, fz;
    uint32_t keep;  // bit c: corner c is inside the volume
};

GARG_DEV TexelCorners fetch_corners(const VolumeView& v, f3 p) {
    TexelCorners c;
    float tx = (p.x + v.half_xy) * v.scale_xy - 0.5f;
    float ty = (p.y + v.half_xy) * v.scale_xy - 0.5f;
    float tz = (p.z + v.half_z) * v.scale_z - 0.5f;
    float bx = floorf(tx), by = floorf(ty), bz = floorf(tz);
    c.fx = tx - bx;
    c.fy = ty - by;
    c.fz = tz - bz;
    int x0 = (int)bx, y0 = (int)by, z0 = (int)bz;
    const __attribute__((address_space(1))) Voxel* g = (const __attribute__((address_space(1))) Voxel*)v.voxels;
    c.keep = 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        int x = x0 + (k & 1), y = y0 + ((k >> 1) & 1), z = z0 + ((k >> 2) & 1);
        bool inside = (unsigned)x < (unsigned)v.size && (unsigned)y < (unsigned)v.size && (unsigned)z < (unsigned)v.size;
        c.keep |= inside ? (1u << k) : 0u;
        uint32_t i = inside ? ((uint32_t)z * (uint32_t)v.size + (uint32_t)y) * (uint32_t)v.size + (uint32_t)x : 0u;
        c.v[k].emission = g[i].emission;
        c.v[k].velocity = g[i].velocity;
    }
    return c;
}

GARG_DEV Sample combine_corners(const TexelCorners& c) {
    Sample s{f3{0, 0, 0}, 0.0f, f3{0, 0, 0}};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float w = ((k & 1) ? c.fx : 1.0f - c.fx) * (((k >> 1) & 1) ? c.fy : 1.0f - c.fy) *
                  (((k >> 2) & 1) ? c.fz : 1.0f - c.fz);
        if (((c.keep >> k) & 1u) && w != 0.0f) add_corner(s, c.v[k], w);
    }
    return s;
}

// Linear lookup in an n entry table at a normalized coordinate with texel
// centres at (i + 0.5) / n, clamped to the edge entries.
GARG_DEV float table_linear(const float* table, int n, float coordinate) {
    float x = coordinate * (float)n - 0.5f;
    float b = floorf(x);
    float f = x - b;
    int i0 = min(max((int)b, 0), n - 1);
    int i1 = min(max((int)b + 1, 0), n - 1);
    return table[i0] * (1.0f - f) + table[i1] * f;
}

GARG_DEV f3 table_linear3(const float4* table, int n, float coordinate) {
    float x = coordinate * (float)n - 0.5f;
    float b = floorf(x);
    float f = x - b;
    int i0 = min(max((int)b, 0), n - 1);
    int i1 = min(max((int)b + 1, 0), n - 1);
    float4 a = table[i0], c = table[i1];
    return f3{a.x * (1.0f - f) + c.x * f, a.y * (1.0f - f) + c.y * f, a.z * (1.0f - f) + c.z * f};
}

// Log-spaced blackbody coordinate (Common.h:56-59).
GARG_DEV float blackbody_coordinate(float temperature) {
    float t = fminf(fmaxf(temperature, GARG_F(BLACKBODY_T_MIN)), GARG_F(BLACKBODY_T_MAX));
    return (logf(t) - logf(GARG_F(BLACKBODY_T_MIN))) / (logf(GARG_F(BLACKBODY_T_MAX)) - logf(GARG_F(BLACKBODY_T_MIN)));
}

// Procedural stars on a six-face cube map of 96 x 96 cells, at most one star
// per cell, drawn from pcg4d (Geodesic.h:190-215).
GARG_DEV f3 starfield(f3 d, uint32_t seed, const float4* blackbody) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    uint32_t face;
    float s0, s1, major;
    if (ax >= ay && ax >= az) {
        face = d.x > 0.0f ? 0u : 1u;
        s0 = d.y;
        s1 = d.z;
        major = ax;
    } else if (ay >= az) {
        face = d.y > 0.0f ? 2u : 3u;
        s0 = d.x;
        s1 = d.z;
        major = ay;
    } else {
        face = d.z > 0.0f ? 4u : 5u;
        s0 = d.x;
        s1 = d.y;
        major = az;
    }
    const float cells = GARG_F(STAR_CELLS);
    float u = (s0 / major * 0.5f + 0.5f) * cells;
    float v = (s1 / major * 0.5f + 0.5f) * cells;
    uint32_t cu = (uint32_t)fminf(u, cells - 1.0f), cv = (uint32_t)fminf(v, cells - 1.0f);
    uint4 h = pcg4d(make_uint4(cu, cv, face, seed));
    if (h.x > STAR_THRESHOLD) return f3{0, 0, 0};
    float du = u - ((float)cu + unit_float(h.y));
    float dv = v - ((float)cv + unit_float(h.z));
    float falloff = expf(-(du * du + dv * dv) * 40.0f);
    float w = unit_float(h.w);
    float brightness = w * w * w * falloff;
    float temperature = 3000.0f + (12000.0f - 3000.0f) * unit_float(h.w ^ h.y);
    return table_linear3(blackbody, (int)BLACKBODY_TABLE_SIZE, blackbody_coordinate(temperature)) * brightness;
}

// tanh x for x >= 0 in the fast march forms: its Taylor series to x^7 below
// 1/8 (the first omitted term is 1.3e-9 of the result there), otherwise
// 1 - 2 / (1 + e^(2x)) with the hardware exp2 and reciprocal (a few
// roundings of the complement, about 1e-6 of tanh at x = 1/8 and less
// above).
GARG_DEV float fast_tanh(float x) {
    if (x < 0.125f) {
        float x2 = x * x;
        float p = fmaf(x2, -5.3968253968253968e-02f, 1.3333333333333333e-01f);
        p = fmaf(x2, p, -3.3333333333333333e-01f);
        return fmaf(x * x2, p, x);
    }
    float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);  // e^(2x)
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}

// Redshift factor g = sqrt(f) / (gamma (1 + d . beta)) with the simulated
// speed compressed below c and the tanh argument capped (Geodesic.h:225-236).
// FAST evaluates it with fast_tanh, hardware square roots and reciprocals:
// g = sqrt(f (1 - beta^2)) / (1 + d . beta).
template <bool FAST = false>
GARG_DEV float redshift_factor(float f, f3 direction, f3 gas) {
    if constexpr (FAST) {
        float magnitude = __builtin_amdgcn_sqrtf(dot(gas, gas));
        float speed = fminf(magnitude, 8.0f);
        float beta = speed > 0.0f ? GARG_F(GAS_SPEED_CEILING) * fast_tanh(speed * (1.0f / GARG_F(GAS_SPEED_CEILING)))
                                  : 0.0f;
        f3 velocity = magnitude > 0.0f ? gas * (beta * __builtin_amdgcn_rcpf(magnitude)) : f3{0, 0, 0};
        return __builtin_amdgcn_sqrtf(fmaxf(f, 0.0f) * (1.0f - beta * beta)) *
               __builtin_amdgcn_rcpf(1.0f + dot(direction, velocity));
    } else {
        float magnitude = length(gas);
        float speed = fminf(magnitude, 8.0f);
        float beta = speed > 0.0f ? GARG_F(GAS_SPEED_CEILING) * tanhf(speed / GARG_F(GAS_SPEED_CEILING)) : 0.0f;
        f3 velocity = magnitude > 0.0f ? gas * (beta / magnitude) : f3{0, 0, 0};
        float gamma = 1.0f / sqrtf(1.0f - beta * beta);
        return sqrtf(fmaxf(f, 0.0f)) / (gamma * (1.0f + dot(direction, velocity)));
    }
}

// Front-to-back compositing state of one ray (march.metal:56-64, 81-102).
struct Gather {
    f3 color;
    float transmittance;
    float travelled;
    float weighted_depth;
    float depth_weight;

    GARG_DEV void reset() {
        color = f3{0, 0, 0};
        transmittance = 1.0f;
        travelled = 0.0f;
        weighted_depth = 0.0f;
        depth_weight = 0.0f;
    }

    // One sample at `p` after a step of length `path`; `f` is the metric
    // factor for the gravitational redshift and `direction_of` yields the
    // ray's locally measured direction when Doppler beaming needs it. FAST
    // (the fast march forms) takes the redshift factor's fast evaluation,
    // the hardware exp2 for the opacity and a product for the beaming
    // weight's division by a constant.
    template <bool FAST = false, class DirectionOf>
    GARG_DEV void sample(const VolumeView& vol, float opacity, bool redshift, f3 p, float path, float f,
                         DirectionOf direction_of) {
        if (!(fabsf(p.z) < GARG_F(DISK_SLAB_HALF_HEIGHT) && fabsf(p.x) < vol.half_xy && fabsf(p.y) < vol.half_xy &&
              transmittance > GARG_F(TRANSMITTANCE_FLOOR)))
            return;
        Sample s = sample_volume(vol, p);
        if (!(s.density > 0.0f)) return;
        float g3 = 1.0f;
        if (redshift) {
            float g = redshift_factor<FAST>(f, direction_of(), s.velocity);
            float gravity = FAST ? __builtin_amdgcn_sqrtf(fmaxf(f, 0.0f)) : sqrtf(fmaxf(f, 0.0f));
            float bulk = fminf(FAST ? s.density * (1.0f / GARG_F(BEAMING_DENSITY_FLOOR))
                                    : s.density / GARG_F(BEAMING_DENSITY_FLOOR),
                               1.0f);
            float gr3 = gravity * gravity * gravity;
            g3 = gr3 + (g * g * g - gr3) * bulk;
        }
        float alpha = FAST ? 1.0f - __builtin_amdgcn_exp2f(-opacity * s.density * path * 1.4426950408889634f)
                           : 1.0f - expf(-opacity * s.density * path);
        f3 contribution = s.emission * (transmittance * g3 * path);
        color = color + contribution;
        float weight = contribution.x * 0.2126f + contribution.y * 0.7152f + contribution.z * 0.0722f;
        weighted_depth += weight * travelled;
        depth_weight += weight;
        transmittance *= 1.0f - alpha;
    }

    GARG_DEV float depth() const { return depth_weight > 0.0f ? weighted_depth / depth_weight : travelled; }
};

// Everything a march kernel reads besides the uniforms.
struct Scene {
    VolumeView volume;          // voxels == nullptr: no gas
    const float* sphere_table;  // SWEEP_TABLE_SIZE entries each
    const float* camera_table;
    const float* sky_table;
    const float4* blackbody;    // BLACKBODY_TABLE_SIZE entries, rgb + 1
};

// A copy of `v` with every 32-bit word read from the first active lane. For
// data the whole wave shares this keeps it in scalar registers: the loop
// that uses it then reads SGPR operands instead of reloading per-lane copies.
template <class T>
GARG_DEV T wave_uniform(const T& v) {
    static_assert(sizeof(T) % 4 == 0, "word-sized copy");
    T out;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&v);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return out;
}

// Sum of a value over the active lanes of a wave64.
GARG_DEV uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
    return v;
}

// Counters of one ray, summed per wave and added with one atomic each.
GARG_DEV void add_counters(garg_march_counters* counters, bool active, bool over, bool exhausted, bool captured,
                           uint32_t steps) {
    uint64_t m_rays = __ballot(active), m_over = __ballot(active && over);
    uint64_t m_exh = __ballot(active && exhausted), m_cap = __ballot(active && captured);
    uint64_t total = wave_sum_u64(active ? (uint64_t)steps : 0ull);
    if (__lane_id() == 0) {
        uint32_t* c = reinterpret_cast<uint32_t*>(counters);
        if (m_rays) atomicAdd(&c[0], (uint32_t)__popcll(m_rays));
        if (m_over) atomicAdd(&c[1], (uint32_t)__popcll(m_over));
        if (m_exh) atomicAdd(&c[2], (uint32_t)__popcll(m_exh));
        if (m_cap) atomicAdd(&c[3], (uint32_t)__popcll(m_cap));
        if (total) atomicAdd((unsigned long long*)&counters->steps, (unsigned long long)total);
    }
}

// Camera ray through pixel centre plus jitter (march.metal:215-218).
GARG_DEV f3 camera_direction(const garg_march_uniforms& u, uint32_t px, uint32_t py) {
    float sx = (float)px + 0.5f + u.jitter[0];
    float sy = (float)py + 0.5f + u.jitter[1];
    float nx = (2.0f * sx / (float)u.resolution[0] - 1.0f) * u.tan_half_fov[0];
    float ny = (1.0f - 2.0f * sy / (float)u.resolution[1]) * u.tan_half_fov[1];
    f3 d = load3(u.camera_forward) + load3(u.camera_right) * nx + load3(u.camera_up) * ny;
    return normalize(d);
}

GARG_DEV void store_half4(h4* p, float a, float b, float c, float d) {
    h4 v;
    v.x = (_Float16)a;
    v.y = (_Float16)b;
    v.z = (_Float16)c;
    v.w = (_Float16)d;
    *p = v;
}

}  // namespace garg
