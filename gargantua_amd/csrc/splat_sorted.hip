// The exact (eight-corner) deposit with the atomics aggregated in LDS.
//
// One no-return 32-bit global atomic costs about as much as the whole
// deposit's arithmetic: the bin is bound by their rate (about 27 G/s on
// gfx950 to random words, which the splat at 48 atomics per particle nearly
// reaches). Particles are therefore radix-sorted by the voxel their low CIC
// corner sits in; a workgroup then deposits a run of sorted particles, which
// covers a few rows of one z plane, into an LDS box with LDS atomics and
// flushes each non-zero box word with one global atomic. The sums are
// integers, so the result is bit-identical to the direct deposit whatever
// the order. A run whose box does not fit (it crosses a z plane, or spans
// too many rows) deposits directly, like splat_kernel.
#include <hipcub/hipcub.hpp>

#include "device_common.hpp"
#include "launch.hpp"

namespace garg {

namespace {

constexpr uint32_t SORTED_BLOCK = 256;
constexpr uint32_t BOX_WORDS = 12288;  // 48 KB of LDS

// Sorted particles per thread: a workgroup's run should cover few enough
// rows of one z plane for its box (rows x size x 6 words) to fit, and a
// finer grid puts fewer particles in a row.
__host__ __device__ inline uint32_t sorted_per_thread(uint32_t size) {
    return size <= 128u ? 8u : size <= 192u ? 4u : size <= 256u ? 2u : 1u;
}

struct Corners {
    int low[3], high[3];
    float fraction[3];
    float energy, relative, vx, vy, vz;
};

// The same float operations as splat_kernel, so the integer contributions
// are the same bits.
GARG_DEV bool corners_of(const SplatArgs& a, uint32_t i, Corners& c, int floor_out[3]) {
    const float half[3] = {a.half_xy, a.half_xy, a.half_z};
    const float pos[3] = {a.p.px[i], a.p.py[i], a.p.pz[i]};
    float cube[3];
    bool outside = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        cube[k] = (pos[k] + half[k]) / (2.0f * half[k]);
        outside = outside || cube[k] < 0.0f || cube[k] >= 1.0f;
    }
    if (outside) return false;
    float relative = a.p.temperature[i] / a.peak_temperature;
    float energy = relative * relative;
    energy = energy * relative;
    energy = energy * relative;
    energy = energy * a.emission_scale;
    c.energy = energy;
    c.relative = relative;
    c.vx = a.p.vx[i];
    c.vy = a.p.vy[i];
    c.vz = a.p.vz[i];
    const int n = (int)a.size;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float grid = cube[k] * (float)a.size;
        grid = grid - 0.5f;
        float base = floorf(grid);
        c.fraction[k] = grid - base;
        floor_out[k] = (int)base;
        c.high[k] = min(floor_out[k] + 1, n - 1);
        c.low[k] = max(floor_out[k], 0);
    }
    return true;
}

struct Contribution {
    uint32_t e, et, mass;
    int32_t mx, my, mz;
};

GARG_DEV Contribution contribution(float weight, const Corners& c) {
    Contribution out;
    float e = c.energy * weight;
    e = e * GARG_F(ENERGY_FIXED);
    e = e + 0.5f;
    float et = c.energy * c.relative;
    et = et * weight;
    et = et * GARG_F(ENERGY_T_FIXED);
    et = et + 0.5f;
    float wm = weight * GARG_F(MOMENTUM_FIXED);
    float m = weight * GARG_F(MASS_FIXED);
    m = m + 0.5f;
    out.e = (uint32_t)e;
    out.et = (uint32_t)et;
    out.mx = (int32_t)rintf(c.vx * wm);
    out.my = (int32_t)rintf(c.vy * wm);
    out.mz = (int32_t)rintf(c.vz * wm);
    out.mass = (uint32_t)m;
    return out;
}

GARG_DEV float corner_weight(const Corners& c, uint32_t k) {
    float wx = (k & 1u) ? c.fraction[0] : 1.0f - c.fraction[0];
    float wy = (k & 2u) ? c.fraction[1] : 1.0f - c.fraction[1];
    float wz = (k & 4u) ? c.fraction[2] : 1.0f - c.fraction[2];
    float w = wx * wy;
    return w * wz;
}

GARG_DEV void global_add_u32(uint32_t* p, uint32_t v) {
    if (v) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
GARG_DEV void global_add_i32(int32_t* p, int32_t v) {
    if (v) __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

// Sort key of a particle: the floor of its voxel-centre grid coordinate plus
// one on each axis, in a (size + 1)^3 index (z major); particles outside the
// cube get (size + 1)^3 and sort last.
__global__ __launch_bounds__(256) void splat_key_kernel(SplatArgs a, uint32_t* keys, uint32_t* order) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.p.count) return;
    const uint32_t m = a.size + 1u;
    Corners c;
    int f[3];
    uint32_t key = m * m * m;
    if (corners_of(a, i, c, f)) key = ((uint32_t)(f[2] + 1) * m + (uint32_t)(f[1] + 1)) * m + (uint32_t)(f[0] + 1);
    keys[i] = key;
    order[i] = i;
}

__global__ __launch_bounds__(SORTED_BLOCK) void splat_sorted_kernel(SplatArgs a, const uint32_t* keys,
                                                                   const uint32_t* order, uint32_t per_thread) {
    __shared__ uint32_t box[BOX_WORDS];
    __shared__ uint32_t s_plan[4];  // z0, y0, rows (0: deposit directly), valid count
    const uint32_t m = a.size + 1u;
    const uint32_t outside = m * m * m;
    const uint32_t first = blockIdx.x * SORTED_BLOCK * per_thread;
    const uint32_t last_index = min(first + SORTED_BLOCK * per_thread, a.p.count) - 1u;
    const uint32_t n = a.size;

    if (threadIdx.x == 0) {
        // The run's low corners lie between its first and last keys; with
        // one z plane they cover rows y0..y1 and their +1 corners y1 + 1 and
        // z0 + 1, clamped to the grid.
        uint32_t k0 = keys[first];
        uint32_t k1 = keys[last_index];
        uint32_t rows = 0, z0 = 0, y0 = 0;
        if (k0 != outside) {
            if (k1 == outside) {
                // Only the valid prefix counts: find its last key.
                uint32_t lo = first, hi = last_index;  // keys[lo] valid, keys[hi] outside
                while (hi - lo > 1u) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (keys[mid] == outside) hi = mid;
                    else lo = mid;
                }
                k1 = keys[lo];
            }
            uint32_t fz0 = k0 / (m * m), fy0 = (k0 / m) % m;
            uint32_t fz1 = k1 / (m * m), fy1 = (k1 / m) % m;
            if (fz0 == fz1) {
                // Low corners at max(f, 0) and high corners at min(f + 1, n - 1)
                // in grid units, where f = key coordinate - 1.
                int ylo = max((int)fy0 - 1, 0), yhi = min((int)fy1, (int)n - 1);
                int zlo = max((int)fz0 - 1, 0), zhi = min((int)fz0, (int)n - 1);
                uint32_t r = (uint32_t)(yhi - ylo + 1) * (uint32_t)(zhi - zlo + 1);
                if ((uint64_t)r * n * 6u <= BOX_WORDS) {
                    rows = r;
                    z0 = (uint32_t)zlo;
                    y0 = (uint32_t)ylo;
                    s_plan[3] = (uint32_t)(yhi - ylo + 1);
                }
            }
        }
        s_plan[0] = z0;
        s_plan[1] = y0;
        s_plan[2] = rows;
    }
    __syncthreads();
    const uint32_t rows = s_plan[2];
    const uint32_t z0 = s_plan[0], y0 = s_plan[1];
    const uint32_t ycount = rows ? s_plan[3] : 1u;
    const uint32_t box_words = rows * n * 6u;
    for (uint32_t w = threadIdx.x; w < box_words; w += SORTED_BLOCK) box[w] = 0u;
    __syncthreads();

    // Each thread deposits `per_thread` consecutive sorted particles,
    // so the lanes of one LDS atomic are that many particles apart and
    // rarely share a voxel (neighbours in the sorted order mostly do).
    for (uint32_t j = 0; j < per_thread; ++j) {
        const uint32_t s = first + threadIdx.x * per_thread + j;
        if (s > last_index || keys[s] == outside) continue;
        const uint32_t i = order[s];
        Corners c;
        int f[3];
        corners_of(a, i, c, f);
#pragma unroll
        for (uint32_t k = 0; k < CIC_CORNERS; ++k) {
            int cx = (k & 1u) ? c.high[0] : c.low[0];
            int cy = (k & 2u) ? c.high[1] : c.low[1];
            int cz = (k & 4u) ? c.high[2] : c.low[2];
            float w = corner_weight(c, k);
            if (w <= 0.0f) continue;
            Contribution v = contribution(w, c);
            if (rows) {
                // Box word of sum q at (x, y, z): ((q * rows + row) * n + x),
                // row = (z - z0) * ycount + (y - y0).
                uint32_t row = ((uint32_t)cz - z0) * ycount + ((uint32_t)cy - y0);
                uint32_t base = row * n + (uint32_t)cx;
                const uint32_t plane = rows * n;
                if (v.e) atomicAdd(&box[base], v.e);
                if (v.et) atomicAdd(&box[plane + base], v.et);
                if (v.mx) atomicAdd(&box[2u * plane + base], (uint32_t)v.mx);
                if (v.my) atomicAdd(&box[3u * plane + base], (uint32_t)v.my);
                if (v.mz) atomicAdd(&box[4u * plane + base], (uint32_t)v.mz);
                if (v.mass) atomicAdd(&box[5u * plane + base], v.mass);
            } else {
                uint32_t index = ((uint32_t)cz * n + (uint32_t)cy) * n + (uint32_t)cx;
                global_add_u32(&a.emission[2u * index + 0u], v.e);
                global_add_u32(&a.emission[2u * index + 1u], v.et);
                global_add_i32(&a.momentum[3u * index + 0u], v.mx);
                global_add_i32(&a.momentum[3u * index + 1u], v.my);
                global_add_i32(&a.momentum[3u * index + 2u], v.mz);
                global_add_u32(&a.mass[index], v.mass);
            }
        }
    }
    if (!rows) return;
    __syncthreads();
    // Flush: one global atomic per non-zero box word.
    const uint32_t plane = rows * n;
    for (uint32_t w = threadIdx.x; w < plane; w += SORTED_BLOCK) {
        uint32_t row = w / n, x = w % n;
        uint32_t z = z0 + row / ycount, y = y0 + row % ycount;
        uint32_t index = (z * n + y) * n + x;
        global_add_u32(&a.emission[2u * index + 0u], box[w]);
        global_add_u32(&a.emission[2u * index + 1u], box[plane + w]);
        global_add_i32(&a.momentum[3u * index + 0u], (int32_t)box[2u * plane + w]);
        global_add_i32(&a.momentum[3u * index + 1u], (int32_t)box[3u * plane + w]);
        global_add_i32(&a.momentum[3u * index + 2u], (int32_t)box[4u * plane + w]);
        global_add_u32(&a.mass[index], box[5u * plane + w]);
    }
}

size_t splat_sort_temp_bytes(uint32_t count) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (const uint32_t*)nullptr, (uint32_t*)nullptr, count, 0, 32);
    return bytes;
}

hipError_t launch_splat_sorted(const SplatArgs& a, const SplatSortBuffers& b, hipStream_t stream) {
    const uint32_t count = a.p.count;
    splat_key_kernel<<<dim3((count + 255u) / 256u), 256, 0, stream>>>(a, b.keys_in, b.order_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint64_t m = (uint64_t)a.size + 1u;
    int bits = 1;
    while ((1ull << bits) <= m * m * m) ++bits;
    size_t temp = b.temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(b.temp, temp, b.keys_in, b.keys_out, b.order_in, b.order_out, count, 0, bits,
                                           stream);
    if (e != hipSuccess) return e;
    const uint32_t per = sorted_per_thread(a.size);
    const uint32_t run = SORTED_BLOCK * per;
    splat_sorted_kernel<<<dim3((count + run - 1u) / run), SORTED_BLOCK, 0, stream>>>(a, b.keys_out, b.order_out, per);
    return hipGetLastError();
}

}  // namespace garg
