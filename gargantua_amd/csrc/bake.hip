// The bake strategy for gfx950 (bake.metal): for a camera that holds still,
// each ray's sample points inside the disk slab are integrated once and
// stored, and every later frame walks the stored samples against the current
// volume. Samples live in structure-of-arrays slabs, sample k of ray i at
// k * rays + i, so the lanes of a wave read and write consecutive 8-byte
// records at every k.
#include "device_march.hpp"
#include "launch.hpp"

// Lanes that share one ray of the walk (1, 2, 4 or 8).
#define GARG_WALK_LANES 4

namespace garg {

GARG_DEV uint32_t pack_half2(float a, float b) {
    _Float16 ha = (_Float16)a, hb = (_Float16)b;
    uint16_t ua = __builtin_bit_cast(uint16_t, ha), ub = __builtin_bit_cast(uint16_t, hb);
    return (uint32_t)ua | ((uint32_t)ub << 16);
}

GARG_DEV float2 unpack_half2(uint32_t v) {
    _Float16 a = __builtin_bit_cast(_Float16, (uint16_t)(v & 0xffffu));
    _Float16 b = __builtin_bit_cast(_Float16, (uint16_t)(v >> 16));
    return make_float2((float)a, (float)b);
}

// bakeRays (bake.metal:17-92): the still-policy Schwarzschild march that
// stores half4(position, path since the last sample) at least bake_spacing
// apart while inside the slab, up to `capacity` samples.
__global__ __launch_bounds__(64) void bake_kernel(BakeArgs a) {
    const garg_march_uniforms& u = a.uniforms;
    uint32_t px = blockIdx.x * 8u + (threadIdx.x & 7u);
    uint32_t py = a.row0 + blockIdx.y * 8u + (threadIdx.x >> 3);
    if (px >= u.resolution[0] || py >= u.resolution[1] || py >= a.row0 + a.rows) return;
    const uint32_t rays = u.resolution[0] * u.resolution[1];
    const uint32_t index = py * u.resolution[0] + px;
    garg_march_uniforms still = u;
    still.jitter[0] = still.jitter[1] = 0.0f;
    f3 dir = camera_direction(still, px, py);
    Prepared p = prepare_ray(load3(u.camera_position), dir, u, a.scene);
    if (p.sky_only) {
        a.headers[index] = make_uint4(0u, 0u, pack_half2(p.sky.x, p.sky.y), pack_half2(p.sky.z, 1.0f));
        return;
    }
    PlaneRay ray = p.ray;
    RayState carry{0, 0, 0, 0, 0};
    f3 prev = ray_position(ray);
    float since = 0.0f, travelled = 0.0f, first_depth = -1.0f;
    uint32_t count = 0, outcome = GARG_RAY_EXHAUSTED;
    bool overflowed = false, left_sphere = false;
    const garg_geodesic_settings& gs = u.geodesic;
    const float hx = u.volume_half_extent;
    for (uint32_t k = 0; k < gs.step_cap; ++k) {
        float h = ray_step_length(prev.z, ray.s.r, gs);
        RayState before = ray.s;
        ray_apply(ray.s, carry, ray_increment(ray.s, h), false);
        if (u.sphere_radius > 0.0f && ray.s.r >= u.sphere_radius && ray.s.rdot > 0.0f && before.r < u.sphere_radius) {
            float hh = h * (u.sphere_radius - before.r) / (ray.s.r - before.r);
            ray.s = before;
            ray_apply(ray.s, carry, ray_increment(before, hh), false);
        }
        f3 pos = ray_position(ray);
        float path = length(pos - prev);
        since += path;
        travelled += path;
        prev = pos;
        if (ray.s.r <= gs.capture_radius) {
            outcome = GARG_RAY_CAPTURED;
            break;
        }
        if (fabsf(pos.z) < GARG_F(DISK_SLAB_HALF_HEIGHT) && fabsf(pos.x) < hx && fabsf(pos.y) < hx &&
            since >= u.bake_spacing) {
            if (count < a.capacity) {
                store_half4(&a.samples[(size_t)count * rays + index], pos.x, pos.y, pos.z, since);
                count += 1u;
                if (first_depth < 0.0f) first_depth = travelled;
            } else {
                overflowed = true;
            }
            since = 0.0f;
        }
        if (u.sphere_radius > 0.0f && ray.s.r >= u.sphere_radius && ray.s.rdot > 0.0f) {
            outcome = GARG_RAY_ESCAPED;
            left_sphere = true;
            break;
        }
        if (ray.s.r >= gs.escape_radius) {
            outcome = GARG_RAY_ESCAPED;
            break;
        }
    }
    bool falls = outcome == GARG_RAY_CAPTURED ||
                 (outcome == GARG_RAY_EXHAUSTED && ray.s.rdot < 0.0f &&
                  fabsf(ray.angular_momentum) < GARG_F(B_CRITICAL) * ray.energy);
    f3 fin = left_sphere ? sphere_exit_direction(ray, u, a.scene) : ray_direction(ray);
    float depth = first_depth >= 0.0f ? first_depth : travelled;
    a.headers[index] = make_uint4(count, falls ? 1u : 0u, pack_half2(fin.x, fin.y),
                                  pack_half2(fin.z, fminf(depth / (2.0f * GARG_F(R_ESCAPE)), 1.0f)));
    if (overflowed) atomicAdd(a.overflow, 1u);
}

// walkRays (bake.metal:94-144): composite the stored samples against the
// current volume with g^3 (no beaming floor, as the reference walk does),
// then the starfield along the stored direction.
//
// The walk of one ray is a chain of dependent loads (a stored point, then its
// eight texels) up to 96 long, and the frame's rays are few enough to be
// resident at once, so the frame takes as long as its longest chains. LANES
// lanes therefore share a ray: lane j of the group loads and shades samples
// j, j + LANES, j + 2 LANES, ..., software-pipelined (while one sample is
// shaded the texels of its next and the point after that are in flight), so
// each chain is LANES times shorter. The composite stays the reference's
// sequential recurrence, in its order: in each round of LANES samples the
// running (transmittance, colour) passes from lane to lane of the group by a
// DPP move and each lane applies its own sample, with the same
// expressions and the same transmittance floor test, so the result is the
// one-lane walk's bit for bit. The sample arithmetic is sample_volume's, in
// the same order. Samples of one round sit at k * rays + i for LANES
// consecutive k, and a wave covers 64 / LANES consecutive pixels of a row.
template <int LANES>
GARG_DEV float from_previous_lane(float v) {
    // Lane j of each group of LANES takes lane j - 1's value (what lane 0
    // receives is not used).
    if constexpr (LANES == 1) return v;
    // quad_perm [0,0,2,2] / [0,0,1,2]; row_shr:1 for groups of 8 inside a row of 16.
    constexpr int ctrl = LANES == 2 ? 0xA0 : LANES == 4 ? 0x90 : 0x111;
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctrl, 0xF, 0xF, false));
}

template <int LANES>
GARG_DEV float from_last_lane(float v) {
    // Every lane of a group of LANES takes the group's last lane's value.
    if constexpr (LANES == 1) return v;
    if constexpr (LANES == 8) {
        // ds_swizzle bit mode: lane (i & 0x18) | 7 of its 32-lane half.
        return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x18 | (0x07 << 5)));
    }
    constexpr int ctrl = LANES == 4 ? 0xFF : 0xF5;  // quad_perm [3,3,3,3] / [1,1,3,3]
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctrl, 0xF, 0xF, false));
}

template <int LANES>
__global__ __launch_bounds__(64) void walk_kernel(WalkArgs a) {
    static_assert(LANES == 1 || LANES == 2 || LANES == 4 || LANES == 8, "a ray's lanes must fit one DPP row");
    const garg_march_uniforms& u = a.uniforms;
    const uint32_t lane = threadIdx.x % LANES;
    const uint32_t px = blockIdx.x * (64u / LANES) + threadIdx.x / LANES;
    const uint32_t py = blockIdx.y;
    // Whole groups leave together: a group is one pixel.
    if (px >= u.resolution[0] || py >= u.resolution[1]) return;
    const uint32_t rays = u.resolution[0] * u.resolution[1];
    const uint32_t index = py * u.resolution[0] + px;
    const uint4 header = a.headers[index];
    const uint32_t count = min(header.x, a.capacity);
    f3 color{0, 0, 0};
    float transmittance = 1.0f;
    f3 before = load3(u.camera_position);  // the point of the sample before the round's first
    const bool gas = a.scene.volume.voxels != nullptr;
    if (gas && count > 0u) {
        const h4* slab = a.samples + index;
        h4 point{}, ahead{};
        TexelCorners texels{};
        if (lane < count) {
            point = slab[(size_t)lane * rays];
            texels = fetch_corners(a.scene.volume, f3{(float)point.x, (float)point.y, (float)point.z});
            if (lane + LANES < count) ahead = slab[(size_t)(lane + LANES) * rays];
        }
        for (uint32_t round = 0; round < count; round += LANES) {
            // The group holds one transmittance, so it leaves together.
            if (transmittance <= GARG_F(TRANSMITTANCE_FLOOR)) break;
            const uint32_t k = round + lane;
            const bool mine = k < count;
            const h4 cur = point;
            // Combine this sample's texels before the next sample's gather is
            // issued, so that one set of eight texels is live per lane.
            Sample s{};
            if (mine) s = combine_corners(texels);
            if (k + LANES < count) {
                point = ahead;
                texels = fetch_corners(a.scene.volume, f3{(float)point.x, (float)point.y, (float)point.z});
                if (k + 2u * LANES < count) ahead = slab[(size_t)(k + 2u * LANES) * rays];
            }
            const f3 pos{(float)cur.x, (float)cur.y, (float)cur.z};
            const float path = (float)cur.w;
            f3 prev{from_previous_lane<LANES>(pos.x), from_previous_lane<LANES>(pos.y),
                    from_previous_lane<LANES>(pos.z)};
            if (lane == 0u) prev = before;
            bool lit = false;
            float g3 = 1.0f, alpha = 0.0f;
            f3 emission{0, 0, 0};
            if (mine) {
                if (s.density > 0.0f) {
                    lit = true;
                    if (u.redshift != 0u) {
                        float g = redshift_factor(schw_f(length(pos)), normalize(pos - prev), s.velocity);
                        g3 = g * g * g;
                    }
                    alpha = 1.0f - expf(-u.opacity_scale * s.density * path);
                    emission = s.emission;
                }
            }
#pragma unroll
            for (uint32_t step = 0; step < (uint32_t)LANES; ++step) {
                if (step > 0u) {
                    const float t_in = from_previous_lane<LANES>(transmittance);
                    const f3 c_in{from_previous_lane<LANES>(color.x), from_previous_lane<LANES>(color.y),
                                  from_previous_lane<LANES>(color.z)};
                    if (lane == step) {
                        transmittance = t_in;
                        color = c_in;
                    }
                }
                if (lane == step && lit && transmittance > GARG_F(TRANSMITTANCE_FLOOR)) {
                    color = color + emission * (transmittance * g3 * path);
                    transmittance *= 1.0f - alpha;
                }
            }
            transmittance = from_last_lane<LANES>(transmittance);
            color = f3{from_last_lane<LANES>(color.x), from_last_lane<LANES>(color.y), from_last_lane<LANES>(color.z)};
            before = f3{from_last_lane<LANES>(pos.x), from_last_lane<LANES>(pos.y), from_last_lane<LANES>(pos.z)};
        }
    }
    if (lane != 0u) return;
    float2 xy = unpack_half2(header.z), zd = unpack_half2(header.w);
    if (header.y == 0u && u.star_brightness > 0.0f) {
        f3 d = normalize(f3{xy.x, xy.y, zd.x});
        color = color + starfield(d, u.star_seed, a.scene.blackbody) * (transmittance * u.star_brightness);
    }
    store_half4(&a.color[index], color.x, color.y, color.z, 1.0f);
    if (a.depth) a.depth[index] = zd.y;
}

hipError_t launch_bake(const BakeArgs& a, hipStream_t stream) {
    dim3 grid((a.uniforms.resolution[0] + 7u) / 8u, (a.rows + 7u) / 8u);
    bake_kernel<<<grid, 64, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_walk(const WalkArgs& a, hipStream_t stream) {
    const uint32_t per_wave = 64u / GARG_WALK_LANES;
    dim3 grid((a.uniforms.resolution[0] + per_wave - 1u) / per_wave, a.uniforms.resolution[1]);
    walk_kernel<GARG_WALK_LANES><<<grid, 64, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace garg
