// Splat pass for gfx950 (splat.metal): cloud-in-cell deposit of particles
// into 32-bit fixed point sums with integer atomics, then a resolve into the
// 16-byte voxels the march samples (emission rgb + density, velocity, as
// halves). Integer addition is associative, so the sums are bit-exact
// against a CPU replay of the same float operations whatever the order the
// atomics land in; the library is built with contraction off for that.
#include "device_common.hpp"
#include "launch.hpp"

namespace garg {

GARG_DEV void add_u32(uint32_t* p, uint32_t v) {
    if (v) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

GARG_DEV void add_i32(int32_t* p, int32_t v) {
    if (v) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Fixed point round to nearest: uint(x + 0.5) and rint (splat.metal:25-37).
GARG_DEV void deposit(const SplatArgs& a, uint32_t index, float weight, float energy, float relative, float vx,
                      float vy, float vz) {
    float e = energy * weight;
    e = e * GARG_F(ENERGY_FIXED);
    e = e + 0.5f;
    float et = energy * relative;
    et = et * weight;
    et = et * GARG_F(ENERGY_T_FIXED);
    et = et + 0.5f;
    float wm = weight * GARG_F(MOMENTUM_FIXED);
    float m = weight * GARG_F(MASS_FIXED);
    m = m + 0.5f;
    add_u32(&a.emission[2u * index + 0u], (uint32_t)e);
    add_u32(&a.emission[2u * index + 1u], (uint32_t)et);
    add_i32(&a.momentum[3u * index + 0u], (int32_t)rintf(vx * wm));
    add_i32(&a.momentum[3u * index + 1u], (int32_t)rintf(vy * wm));
    add_i32(&a.momentum[3u * index + 2u], (int32_t)rintf(vz * wm));
    add_u32(&a.mass[index], (uint32_t)m);
}

// Cloud in cell in the space of voxel centres (splat.metal:44-88).
__global__ __launch_bounds__(256) void splat_kernel(SplatArgs a) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.p.count) return;
    const float half[3] = {a.half_xy, a.half_xy, a.half_z};
    const float pos[3] = {a.p.px[i], a.p.py[i], a.p.pz[i]};
    float cube[3];
    bool outside = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        cube[k] = (pos[k] + half[k]) / (2.0f * half[k]);
        outside = outside || cube[k] < 0.0f || cube[k] >= 1.0f;
    }
    if (outside) return;
    float relative = a.p.temperature[i] / a.peak_temperature;
    float energy = relative * relative;
    energy = energy * relative;
    energy = energy * relative;
    energy = energy * a.emission_scale;
    float vx = a.p.vx[i], vy = a.p.vy[i], vz = a.p.vz[i];
    const int n = (int)a.size;
    int low[3], high[3];
    float fraction[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float grid = cube[k] * (float)a.size;
        grid = grid - 0.5f;
        float base = floorf(grid);
        fraction[k] = grid - base;
        low[k] = (int)base;
        high[k] = min(low[k] + 1, n - 1);
        low[k] = max(low[k], 0);
    }
    if (a.exact) {
#pragma unroll
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
            uint32_t index = ((uint32_t)cz * a.size + (uint32_t)cy) * a.size + (uint32_t)cx;
            deposit(a, index, w, energy, relative, vx, vy, vz);
        }
    } else {
        // One corner with the trilinear weights as probabilities: unbiased.
        const uint32_t frame = a.frame_counter ? *a.frame_counter : a.frame;
        uint4 h = pcg4d(make_uint4(i, frame, a.seed_lo, a.seed_hi));
        int cx = unit_float(h.x) < fraction[0] ? high[0] : low[0];
        int cy = unit_float(h.y) < fraction[1] ? high[1] : low[1];
        int cz = unit_float(h.z) < fraction[2] ? high[2] : low[2];
        uint32_t index = ((uint32_t)cz * a.size + (uint32_t)cy) * a.size + (uint32_t)cx;
        deposit(a, index, 1.0f, energy, relative, vx, vy, vz);
    }
}

// Sums to emission per unit volume (tint from the mean temperature) with the
// density in alpha, and mean velocity (splat.metal:93-120). With `clear`
// the sums are zeroed in the same pass, which replaces the separate fill.
__global__ __launch_bounds__(256) void resolve_kernel(ResolveArgs a) {
    uint32_t voxels = a.size * a.size * a.size;
    uint32_t index = blockIdx.x * 256u + threadIdx.x;
    if (index >= voxels) return;
    float dx = 2.0f * a.half_xy / (float)a.size;
    float dz = 2.0f * a.half_z / (float)a.size;
    float inverse_volume = 1.0f / (dx * dx * dz);
    uint2 em = reinterpret_cast<const uint2*>(a.emission)[index];
    int32_t mx = a.momentum[3u * index + 0u], my = a.momentum[3u * index + 1u], mz = a.momentum[3u * index + 2u];
    uint32_t ms = a.mass[index];
    float count = (float)ms / GARG_F(MASS_FIXED);
    float energy = (float)em.x / GARG_F(ENERGY_FIXED);
    float energy_t = (float)em.y / GARG_F(ENERGY_T_FIXED);
    f3 tint{0, 0, 0};
    if (energy > 0.0f) {
        float mean_temperature = energy_t / energy * a.peak_temperature;
        tint = table_linear3(a.blackbody, (int)BLACKBODY_TABLE_SIZE, blackbody_coordinate(mean_temperature));
    }
    float density = count * inverse_volume;
    f3 m{(float)mx / GARG_F(MOMENTUM_FIXED), (float)my / GARG_F(MOMENTUM_FIXED), (float)mz / GARG_F(MOMENTUM_FIXED)};
    f3 v = count > 0.0f ? m / count : f3{0, 0, 0};
    f3 e = tint * (energy * inverse_volume);
    Voxel out;
    out.emission = h4{(_Float16)e.x, (_Float16)e.y, (_Float16)e.z, (_Float16)density};
    out.velocity = h4{(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)1.0f};
    a.voxels[index] = out;
    // Sums already zero need no clearing: most voxels of a disk frame stay
    // empty, so this writes 24 bytes only for the ones a deposit reached.
    if (a.clear && (em.x | em.y | (uint32_t)mx | (uint32_t)my | (uint32_t)mz | ms) != 0u) {
        reinterpret_cast<uint2*>(a.emission)[index] = make_uint2(0u, 0u);
        a.momentum[3u * index + 0u] = 0;
        a.momentum[3u * index + 1u] = 0;
        a.momentum[3u * index + 2u] = 0;
        a.mass[index] = 0u;
    }
}

__global__ __launch_bounds__(256) void fill_zero_kernel(uint4* p, uint64_t n16, uint8_t* tail, uint32_t tail_bytes) {
    uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (; i < n16; i += stride) p[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && threadIdx.x < tail_bytes) tail[threadIdx.x] = 0u;
}

// The atomic ceiling the splat is measured against: every thread adds 1 to
// `adds` words of a `words`-word buffer, chosen by pcg4d(thread, k) — the
// same no-return 32-bit atomics as the deposit, spread over the footprint
// of the sums with no more contention than chance gives.
__global__ __launch_bounds__(256) void atomic_rate_kernel(uint32_t* buffer, uint32_t words, uint32_t adds,
                                                          uint32_t threads) {
    uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= threads) return;
    for (uint32_t k = 0; k < adds; ++k) {
        uint4 h = pcg4d(make_uint4(t, k, 0x9e3779b9u, 0x7f4a7c15u));
        __hip_atomic_fetch_add(&buffer[h.x % words], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

hipError_t launch_atomic_rate(uint32_t* buffer, uint32_t words, uint32_t adds, uint32_t threads, hipStream_t stream) {
    atomic_rate_kernel<<<dim3((threads + 255u) / 256u), 256, 0, stream>>>(buffer, words, adds, threads);
    return hipGetLastError();
}

hipError_t launch_splat_bin(const SplatArgs& a, hipStream_t stream) {
    splat_kernel<<<dim3((a.p.count + 255u) / 256u), 256, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_resolve(const ResolveArgs& a, hipStream_t stream) {
    uint32_t voxels = a.size * a.size * a.size;
    resolve_kernel<<<dim3((voxels + 255u) / 256u), 256, 0, stream>>>(a);
    return hipGetLastError();
}

// Zero fill as a kernel rather than a memset node, so it stays correct
// inside captured graphs. `ptr` must be 16-byte aligned.
hipError_t launch_fill_zero(void* ptr, uint64_t bytes, hipStream_t stream) {
    uint64_t n16 = bytes / 16u;
    uint32_t tail = (uint32_t)(bytes - n16 * 16u);
    uint64_t blocks = (n16 + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    if (blocks == 0) blocks = 1;
    fill_zero_kernel<<<dim3((uint32_t)blocks), 256, 0, stream>>>(static_cast<uint4*>(ptr), n16,
                                                                   static_cast<uint8_t*>(ptr) + n16 * 16u, tail);
    return hipGetLastError();
}

}  // namespace garg
