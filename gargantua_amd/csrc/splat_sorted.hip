// The exact (eight-corner) deposit without global atomics.
//
// Device-scope atomics on gfx950 are performed past the per-XCD L2 (every
// one reaches the memory side), about 27 G/s to random words; the direct
// deposit issues 48 per particle and is bound by them. Here particles are
// radix-sorted (onesweep) by the CIC cell their low corner sits in, with a
// 32-byte record of each particle's corner data carried into sorted order.
// The voxels are then split into bricks, each owned by one workgroup: it
// reads the sorted runs of the cells whose corners can land in the brick,
// deposits those corners into an LDS box with LDS atomics, and adds the box
// to the brick's sums with plain read-modify-writes, since no other
// workgroup writes them. The per-corner contributions are the same float
// operations as the direct deposit (this file is built with contraction
// off) and the sums are integers added modulo 2^32, so the result is
// bit-identical to the direct deposit's, whatever the order.
//
// Cells are indexed by the floor of the voxel-centre grid coordinate plus
// one on each axis, c in [0, size], in a (size + 1)^3 table (z major). A
// cell c deposits on each axis to max(c - 1, 0) (weight 1 - fraction) and
// min(c, size - 1) (weight fraction), so voxel v draws from cells v and
// v + 1 only.
#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>

#include "device_common.hpp"
#include "launch.hpp"

namespace garg {

namespace {

// Threads of a brick's workgroup, and the consecutive sorted particles one
// thread deposits. The brick's shape (voxels per axis, one workgroup per
// brick, its sums held in LDS: 24 bytes per voxel) is a launch choice from
// kSplatBricks; every shape gives the same bits.
constexpr uint32_t BRICK_THREADS = 512;
constexpr uint32_t CHUNK = 8;

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

// splat_kernel's deposit, returning the integers instead of adding them.
GARG_DEV Contribution contribution(float weight, float energy, float relative, float vx, float vy, float vz) {
    Contribution out;
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
    out.e = (uint32_t)e;
    out.et = (uint32_t)et;
    out.mx = (int32_t)rintf(vx * wm);
    out.my = (int32_t)rintf(vy * wm);
    out.mz = (int32_t)rintf(vz * wm);
    out.mass = (uint32_t)m;
    return out;
}

// splat_kernel's corner weight: (wx * wy) * wz.
GARG_DEV float corner_weight(float fx, float fy, float fz, uint32_t k) {
    float wx = (k & 1u) ? fx : 1.0f - fx;
    float wy = (k & 2u) ? fy : 1.0f - fy;
    float wz = (k & 4u) ? fz : 1.0f - fz;
    float w = wx * wy;
    return w * wz;
}

GARG_DEV uint32_t brick_of(uint32_t x, uint32_t y, uint32_t z, const uint32_t shape[3], uint32_t nbx, uint32_t nby) {
    return ((z / shape[2]) * nby + y / shape[1]) * nbx + x / shape[0];
}

}  // namespace

// Sort key of a particle: its cell, z major; particles outside the cube get
// (size + 1)^3 and sort last. The particle's staged record (fractions x y
// z, energy, temperature ratio, velocity x y z) is written in particle
// order here, where the reads are coalesced.
__global__ __launch_bounds__(256) void splat_key_kernel(SplatArgs a, SplatSortBuffers b) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.p.count) return;
    const uint32_t m = a.size + 1u;
    Corners c;
    int f[3];
    uint32_t key = m * m * m;
    if (corners_of(a, i, c, f)) {
        key = ((uint32_t)(f[2] + 1) * m + (uint32_t)(f[1] + 1)) * m + (uint32_t)(f[0] + 1);
        b.records[2u * i + 0u] = make_float4(c.fraction[0], c.fraction[1], c.fraction[2], c.energy);
        b.records[2u * i + 1u] = make_float4(c.relative, c.vx, c.vy, c.vz);
    }
    b.keys_in[i] = key;
    b.order_in[i] = i;
}

// Move each particle's record into sorted order; the first and last
// particles of each cell record the cell's range, and the first stamps the
// bricks of the voxels the cell deposits into.
__global__ __launch_bounds__(256) void splat_stage_kernel(SplatArgs a, SplatSortBuffers b, SplatGather g) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    const uint32_t count = a.p.count;
    if (s >= count) return;
    const uint32_t n = a.size, m = n + 1u;
    const uint32_t key = b.keys_out[s];
    if (key == m * m * m) return;
    const uint32_t i = b.order_out[s];
    b.sorted[2u * s + 0u] = b.records[2u * i + 0u];
    b.sorted[2u * s + 1u] = b.records[2u * i + 1u];
    if (s + 1u == count || b.keys_out[s + 1u] != key) g.ends[key] = s + 1u;
    if (s > 0 && b.keys_out[s - 1u] == key) return;
    g.cells[key] = s;
    const int cx = (int)(key % m), cy = (int)((key / m) % m), cz = (int)(key / (m * m));
    const uint32_t nbx = (n + g.brick[0] - 1) / g.brick[0], nby = (n + g.brick[1] - 1) / g.brick[1];
    // The cell's corners (equal on a clamped axis: the repeated store is
    // harmless).
    for (uint32_t k = 0; k < CIC_CORNERS; ++k) {
        uint32_t x = (uint32_t)((k & 1u) ? min(cx, (int)n - 1) : max(cx - 1, 0));
        uint32_t y = (uint32_t)((k & 2u) ? min(cy, (int)n - 1) : max(cy - 1, 0));
        uint32_t z = (uint32_t)((k & 4u) ? min(cz, (int)n - 1) : max(cz - 1, 0));
        g.bricks[brick_of(x, y, z, g.brick, nbx, nby)] = g.stamp;
    }
}

// One workgroup per stamped brick of BRICK_X x BRICK_Y x BRICK_Z voxels. The
// cells whose corners can land in it are BRICK_X + 1 consecutive keys on
// each of (BRICK_Y + 1) x (BRICK_Z + 1) rows, so their particles are that
// many contiguous runs of the sorted order. The workgroup deposits the
// corners that land in the brick into an LDS box and then adds the box to
// the brick's sums: it is the only writer of those voxels in the launch.
template <uint32_t BRICK_X, uint32_t BRICK_Y, uint32_t BRICK_Z>
__global__ __launch_bounds__(BRICK_THREADS) void splat_brick_kernel(SplatArgs a, SplatSortBuffers b, SplatGather g) {
    constexpr uint32_t VOXELS = BRICK_X * BRICK_Y * BRICK_Z;
    constexpr uint32_t ROWS = (BRICK_Y + 1) * (BRICK_Z + 1);
    constexpr uint32_t CELLS = (BRICK_X + 1) * ROWS;
    __shared__ uint32_t box[6 * VOXELS];
    __shared__ uint32_t row_lo[ROWS], row_hi[ROWS], prefix[ROWS + 1];
    if (g.bricks[blockIdx.x] != g.stamp) return;
    const uint32_t n = a.size, m = n + 1u;
    const uint32_t nbx = (n + BRICK_X - 1) / BRICK_X, nby = (n + BRICK_Y - 1) / BRICK_Y;
    const uint32_t x0 = (blockIdx.x % nbx) * BRICK_X, y0 = ((blockIdx.x / nbx) % nby) * BRICK_Y;
    const uint32_t z0 = (blockIdx.x / (nbx * nby)) * BRICK_Z;
    const uint32_t count = a.p.count;
    for (uint32_t w = threadIdx.x; w < 6 * VOXELS; w += BRICK_THREADS) box[w] = 0u;
    for (uint32_t r = threadIdx.x; r < ROWS; r += BRICK_THREADS) {
        row_lo[r] = 0xffffffffu;
        row_hi[r] = 0u;
    }
    __syncthreads();
    // The sorted range of each row: from the first occupied cell's start to
    // the last one's end. A start or end is trusted only when the sorted
    // keys confirm it, so the tables need no clearing between deposits.
    for (uint32_t c = threadIdx.x; c < CELLS; c += BRICK_THREADS) {
        const uint32_t cx = x0 + c % (BRICK_X + 1), row = c / (BRICK_X + 1);
        const uint32_t cy = y0 + row % (BRICK_Y + 1), cz = z0 + row / (BRICK_Y + 1);
        if (cx > n || cy > n || cz > n) continue;
        const uint32_t cell = (cz * m + cy) * m + cx;
        const uint32_t s = g.cells[cell];
        if (s >= count || b.keys_out[s] != cell || (s > 0 && b.keys_out[s - 1u] == cell)) continue;
        atomicMin(&row_lo[row], s);
        atomicMax(&row_hi[row], g.ends[cell]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t r = 0; r < ROWS; ++r) {
            prefix[r] = total;
            if (row_hi[r] > row_lo[r]) total += row_hi[r] - row_lo[r];
        }
        prefix[ROWS] = total;
    }
    __syncthreads();
    const uint32_t total = prefix[ROWS];
    // Each thread takes CHUNK consecutive particles of the rows (a cell's
    // particles are consecutive) and sums each cell's eight corners in
    // registers, adding them to the box once per cell: LDS atomics to one
    // address from many lanes serialise, and a cell holds several particles.
    for (uint32_t f0 = threadIdx.x * CHUNK; f0 < total; f0 += BRICK_THREADS * CHUNK) {
        const uint32_t f1 = min(f0 + CHUNK, total);
        uint32_t row = 0, hi = ROWS;  // the row r with prefix[r] <= f0 < prefix[r + 1]
        while (hi - row > 1u) {
            const uint32_t mid = (row + hi) >> 1;
            if (prefix[mid] <= f0) row = mid;
            else hi = mid;
        }
        uint32_t s = row_lo[row] + (f0 - prefix[row]);
        uint32_t key = 0xffffffffu;
        uint32_t inside = 0, at[CIC_CORNERS];
        uint32_t acc[CIC_CORNERS][6];
        // The next particle's key and record are loaded one step ahead, so
        // their latency overlaps the current particle's arithmetic.
        uint32_t next_key = b.keys_out[s];
        float4 next0 = b.sorted[2u * s + 0u], next1 = b.sorted[2u * s + 1u];
        for (uint32_t f = f0; f < f1; ++f) {
            const uint32_t cell = next_key;
            const float4 r0 = next0, r1 = next1;
            // Step to the particle after: the same row, or the start of the
            // next non-empty one.
            ++s;
            if (f + 1u < f1) {
                if (f + 1u == prefix[row + 1u]) {
                    do {
                        ++row;
                    } while (prefix[row + 1u] == prefix[row]);
                    s = row_lo[row];
                }
                next_key = b.keys_out[s];
                next0 = b.sorted[2u * s + 0u];
                next1 = b.sorted[2u * s + 1u];
            }
            if (cell != key) {
                // Flush the finished cell, then set up this one.
                if (key != 0xffffffffu) {
#pragma unroll
                    for (uint32_t k = 0; k < CIC_CORNERS; ++k) {
                        if (!((inside >> k) & 1u)) continue;
#pragma unroll
                        for (uint32_t q = 0; q < 6; ++q)
                            if (acc[k][q]) atomicAdd(&box[q * VOXELS + at[k]], acc[k][q]);
                    }
                }
                key = cell;
                const int cx = (int)(key % m), cy = (int)((key / m) % m), cz = (int)(key / (m * m));
                inside = 0;
#pragma unroll
                for (uint32_t k = 0; k < CIC_CORNERS; ++k) {
                    const int vx = (k & 1u) ? min(cx, (int)n - 1) : max(cx - 1, 0);
                    const int vy = (k & 2u) ? min(cy, (int)n - 1) : max(cy - 1, 0);
                    const int vz = (k & 4u) ? min(cz, (int)n - 1) : max(cz - 1, 0);
                    const uint32_t lx = (uint32_t)(vx - (int)x0), ly = (uint32_t)(vy - (int)y0);
                    const uint32_t lz = (uint32_t)(vz - (int)z0);
                    at[k] = (lz * BRICK_Y + ly) * BRICK_X + lx;
                    if (lx < BRICK_X && ly < BRICK_Y && lz < BRICK_Z) inside |= 1u << k;
#pragma unroll
                    for (uint32_t q = 0; q < 6; ++q) acc[k][q] = 0u;
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < CIC_CORNERS; ++k) {
                if (!((inside >> k) & 1u)) continue;
                const float w = corner_weight(r0.x, r0.y, r0.z, k);
                if (w <= 0.0f) continue;
                const Contribution v = contribution(w, r0.w, r1.x, r1.y, r1.z, r1.w);
                acc[k][0] += v.e;
                acc[k][1] += v.et;
                acc[k][2] += (uint32_t)v.mx;
                acc[k][3] += (uint32_t)v.my;
                acc[k][4] += (uint32_t)v.mz;
                acc[k][5] += v.mass;
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < CIC_CORNERS; ++k) {
            if (!((inside >> k) & 1u)) continue;
#pragma unroll
            for (uint32_t q = 0; q < 6; ++q)
                if (acc[k][q]) atomicAdd(&box[q * VOXELS + at[k]], acc[k][q]);
        }
    }
    __syncthreads();
    for (uint32_t at = threadIdx.x; at < VOXELS; at += BRICK_THREADS) {
        const uint32_t x = x0 + at % BRICK_X, y = y0 + (at / BRICK_X) % BRICK_Y, z = z0 + at / (BRICK_X * BRICK_Y);
        if (x >= n || y >= n || z >= n) continue;
        const uint32_t e = box[at], et = box[VOXELS + at], mx = box[2 * VOXELS + at], my = box[3 * VOXELS + at];
        const uint32_t mz = box[4 * VOXELS + at], mass = box[5 * VOXELS + at];
        if ((e | et | mx | my | mz | mass) == 0u) continue;
        const uint32_t index = (z * n + y) * n + x;
        uint32_t* em = a.emission + 2u * index;
        uint32_t* mo = reinterpret_cast<uint32_t*>(a.momentum) + 3u * index;
        em[0] += e;
        em[1] += et;
        mo[0] += mx;
        mo[1] += my;
        mo[2] += mz;
        a.mass[index] += mass;
    }
}

// ---------------------------------------------------------------------------
// The one-corner (stochastic) deposit by a counting sort over voxel bricks.
//
// Each particle lands in one voxel, so the particles are grouped by the
// 16 x 16 x 8 brick of that voxel with a counting sort of the project's own:
// every workgroup of the count pass histograms its chunk of particles by
// brick in LDS, a column scan turns the histograms into each workgroup's
// offset within each brick, and the scatter pass writes each particle's
// integer contribution (a 32-byte record) at its brick's next free slot.
// One workgroup per brick then sums its records in an LDS box and adds the
// box to the brick's sums, as the exact deposit's gather does. Every buffer
// is written in full before it is read, so the four kernels hold no memset
// and replay inside captured graphs. Integer sums modulo 2^32 make the
// result bit-identical to splat_kernel's atomics whatever the order.

namespace {

constexpr uint32_t BIN_X = kSplatBinBrick[0], BIN_Y = kSplatBinBrick[1], BIN_Z = kSplatBinBrick[2];
constexpr uint32_t BIN_VOXELS = BIN_X * BIN_Y * BIN_Z;
constexpr uint32_t BIN_THREADS = 1024;
constexpr uint32_t BIN_SEGMENTS = 16;  // the column scan splits the workgroups' column in this many runs
constexpr uint32_t SCAN_BRICKS = BIN_THREADS / BIN_SEGMENTS;  // bricks per column-scan workgroup
constexpr uint32_t BIN_BRICK_THREADS = 512;
constexpr uint32_t BIN_BATCH = 8;  // particles a thread loads at once in the count and scatter passes
static_assert(kSplatBinGroups <= BIN_SEGMENTS * 16u, "the column scan holds 16 counts per thread");
static_assert(kSplatBinMaxBricks <= BIN_THREADS * 16u, "the brick scan holds 16 bricks per thread");

// The voxel splat_kernel's one-corner deposit picks for particle i at `pos`: the same
// float operations (this file is built with contraction off) and the same
// pcg4d draw, so it is the same voxel.
GARG_DEV bool stochastic_voxel(const SplatArgs& a, uint32_t i, const float pos[3], uint32_t voxel[3]) {
    const float half[3] = {a.half_xy, a.half_xy, a.half_z};
    float cube[3];
    bool outside = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        cube[k] = (pos[k] + half[k]) / (2.0f * half[k]);
        outside = outside || cube[k] < 0.0f || cube[k] >= 1.0f;
    }
    if (outside) return false;
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
    const uint32_t frame = a.frame_counter ? *a.frame_counter : a.frame;
    uint4 h = pcg4d(make_uint4(i, frame, a.seed_lo, a.seed_hi));
    voxel[0] = (uint32_t)(unit_float(h.x) < fraction[0] ? high[0] : low[0]);
    voxel[1] = (uint32_t)(unit_float(h.y) < fraction[1] ? high[1] : low[1]);
    voxel[2] = (uint32_t)(unit_float(h.z) < fraction[2] ? high[2] : low[2]);
    return true;
}

}  // namespace

// Count pass: workgroup w takes particles [w * chunk, (w + 1) * chunk), keys
// each by brick and voxel within the brick (~0 outside the cube) and writes
// its brick histogram as row w of the counts. A thread loads the positions
// of BIN_BATCH particles before it works on any, so their latencies overlap.
__global__ __launch_bounds__(BIN_THREADS) void splat_bin_count_kernel(SplatArgs a, SplatBins g) {
    __shared__ uint32_t hist[kSplatBinMaxBricks];
    for (uint32_t b = threadIdx.x; b < g.bricks; b += BIN_THREADS) hist[b] = 0u;
    __syncthreads();
    const uint32_t begin = blockIdx.x * g.chunk;
    const uint32_t end = min(begin + g.chunk, a.p.count);
    for (uint32_t i0 = begin + threadIdx.x; i0 < end; i0 += BIN_THREADS * BIN_BATCH) {
        float pos[BIN_BATCH][3];
#pragma unroll
        for (uint32_t j = 0; j < BIN_BATCH; ++j) {
            const uint32_t i = i0 + j * BIN_THREADS;
            if (i < end) {
                pos[j][0] = a.p.px[i];
                pos[j][1] = a.p.py[i];
                pos[j][2] = a.p.pz[i];
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < BIN_BATCH; ++j) {
            const uint32_t i = i0 + j * BIN_THREADS;
            if (i >= end) break;
            uint32_t v[3], key = 0xffffffffu;
            if (stochastic_voxel(a, i, pos[j], v)) {
                const uint32_t brick = ((v[2] / BIN_Z) * g.nby + v[1] / BIN_Y) * g.nbx + v[0] / BIN_X;
                key = brick * BIN_VOXELS + ((v[2] % BIN_Z) * BIN_Y + v[1] % BIN_Y) * BIN_X + v[0] % BIN_X;
                atomicAdd(&hist[brick], 1u);
            }
            g.keys[i] = key;
        }
    }
    __syncthreads();
    uint32_t* row = g.counts + (size_t)blockIdx.x * g.bricks;
    for (uint32_t b = threadIdx.x; b < g.bricks; b += BIN_THREADS) row[b] = hist[b];
}

// Column scan: for SCAN_BRICKS bricks per workgroup, each brick's counts
// down the workgroups become exclusive offsets (in place), in BIN_SEGMENTS
// runs of at most 16 workgroups summed in parallel; the brick's total goes
// to totals.
__global__ __launch_bounds__(BIN_THREADS) void splat_bin_scan_kernel(SplatBins g) {
    __shared__ uint32_t run_sum[BIN_SEGMENTS][SCAN_BRICKS];
    const uint32_t lane = threadIdx.x % SCAN_BRICKS, run = threadIdx.x / SCAN_BRICKS;
    const uint32_t b = blockIdx.x * SCAN_BRICKS + lane;
    const uint32_t per = (g.groups + BIN_SEGMENTS - 1u) / BIN_SEGMENTS;
    const uint32_t w0 = run * per;
    uint32_t v[16], sum = 0u;
#pragma unroll
    for (uint32_t j = 0; j < 16u; ++j) {
        const uint32_t w = w0 + j;
        v[j] = (b < g.bricks && j < per && w < g.groups) ? g.counts[(size_t)w * g.bricks + b] : 0u;
        sum += v[j];
    }
    run_sum[run][lane] = sum;
    __syncthreads();
    if (b >= g.bricks) return;
    uint32_t offset = 0u;
    for (uint32_t r = 0; r < run; ++r) offset += run_sum[r][lane];
#pragma unroll
    for (uint32_t j = 0; j < 16u; ++j) {
        const uint32_t w = w0 + j;
        if (j < per && w < g.groups) {
            g.counts[(size_t)w * g.bricks + b] = offset;
            offset += v[j];
        }
    }
    if (run == BIN_SEGMENTS - 1u) g.totals[b] = offset;
}

// Scatter pass: the bricks' starts are the exclusive scan of their totals
// (every workgroup scans them; the first writes them out), each brick's
// cursor starts at its start plus this workgroup's offset, and each particle
// of the chunk takes its brick's next slot for its record: the voxel within
// the brick and splat_kernel's six integers at weight one. As in the count
// pass, a thread loads BIN_BATCH particles before it scatters any.
__global__ __launch_bounds__(BIN_THREADS) void splat_bin_scatter_kernel(SplatArgs a, SplatBins g) {
    __shared__ uint32_t cursor[kSplatBinMaxBricks];
    const uint32_t per = (g.bricks + BIN_THREADS - 1u) / BIN_THREADS;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t mine = 0u;
    for (uint32_t j = 0; j < per; ++j)
        if (b0 + j < g.bricks) mine += g.totals[b0 + j];
    // Exclusive scan of `mine` over the workgroup: within each wave by
    // shuffles, then over the waves' totals (held in cursor[0..15] for the
    // moment).
    const uint32_t lane = threadIdx.x % 64u, wave = threadIdx.x / 64u;
    uint32_t inclusive = mine;
#pragma unroll
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t up = __shfl_up(inclusive, d, 64);
        if (lane >= d) inclusive += up;
    }
    if (lane == 63u) cursor[wave] = inclusive;
    __syncthreads();
    uint32_t start = inclusive - mine;
    for (uint32_t w = 0; w < wave; ++w) start += cursor[w];
    __syncthreads();
    const uint32_t* offsets = g.counts + (size_t)blockIdx.x * g.bricks;
    for (uint32_t j = 0; j < per; ++j) {
        const uint32_t b = b0 + j;
        if (b >= g.bricks) break;
        if (blockIdx.x == 0) g.starts[b] = start;
        cursor[b] = start + offsets[b];
        start += g.totals[b];
    }
    if (blockIdx.x == 0 && threadIdx.x == BIN_THREADS - 1u) g.starts[g.bricks] = start;
    __syncthreads();
    const uint32_t begin = blockIdx.x * g.chunk;
    const uint32_t end = min(begin + g.chunk, a.p.count);
    for (uint32_t i0 = begin + threadIdx.x; i0 < end; i0 += BIN_THREADS * BIN_BATCH) {
        uint32_t key[BIN_BATCH];
        float t[BIN_BATCH], vx[BIN_BATCH], vy[BIN_BATCH], vz[BIN_BATCH];
#pragma unroll
        for (uint32_t j = 0; j < BIN_BATCH; ++j) {
            const uint32_t i = i0 + j * BIN_THREADS;
            key[j] = 0xffffffffu;
            if (i < end) {
                key[j] = g.keys[i];
                t[j] = a.p.temperature[i];
                vx[j] = a.p.vx[i];
                vy[j] = a.p.vy[i];
                vz[j] = a.p.vz[i];
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < BIN_BATCH; ++j) {
            if (key[j] == 0xffffffffu) continue;
            const uint32_t slot = atomicAdd(&cursor[key[j] / BIN_VOXELS], 1u);
            float relative = t[j] / a.peak_temperature;
            float energy = relative * relative;
            energy = energy * relative;
            energy = energy * relative;
            energy = energy * a.emission_scale;
            const Contribution c = contribution(1.0f, energy, relative, vx[j], vy[j], vz[j]);
            g.records[2u * slot + 0u] = make_uint4(key[j] % BIN_VOXELS, c.e, c.et, (uint32_t)c.mx);
            g.records[2u * slot + 1u] = make_uint4((uint32_t)c.my, (uint32_t)c.mz, c.mass, 0u);
        }
    }
}

// One workgroup per brick: its records are one contiguous run; their
// integers go into an LDS box with LDS atomics, and the box is added to the
// brick's sums (this workgroup is their only writer in the launch).
__global__ __launch_bounds__(BIN_BRICK_THREADS) void splat_bin_brick_kernel(SplatArgs a, SplatBins g) {
    __shared__ uint32_t box[6 * BIN_VOXELS];
    const uint32_t brick = blockIdx.x;
    const uint32_t s0 = g.starts[brick], s1 = g.starts[brick + 1u];
    if (s0 == s1) return;
    for (uint32_t w = threadIdx.x; w < 6 * BIN_VOXELS; w += BIN_BRICK_THREADS) box[w] = 0u;
    __syncthreads();
    for (uint32_t s = s0 + threadIdx.x; s < s1; s += BIN_BRICK_THREADS) {
        const uint4 r0 = g.records[2u * s + 0u], r1 = g.records[2u * s + 1u];
        const uint32_t at = r0.x;
        if (r0.y) atomicAdd(&box[at], r0.y);
        if (r0.z) atomicAdd(&box[BIN_VOXELS + at], r0.z);
        if (r0.w) atomicAdd(&box[2 * BIN_VOXELS + at], r0.w);
        if (r1.x) atomicAdd(&box[3 * BIN_VOXELS + at], r1.x);
        if (r1.y) atomicAdd(&box[4 * BIN_VOXELS + at], r1.y);
        if (r1.z) atomicAdd(&box[5 * BIN_VOXELS + at], r1.z);
    }
    __syncthreads();
    const uint32_t n = a.size;
    const uint32_t x0 = (brick % g.nbx) * BIN_X, y0 = ((brick / g.nbx) % g.nby) * BIN_Y;
    const uint32_t z0 = (brick / (g.nbx * g.nby)) * BIN_Z;
    for (uint32_t at = threadIdx.x; at < BIN_VOXELS; at += BIN_BRICK_THREADS) {
        const uint32_t x = x0 + at % BIN_X, y = y0 + (at / BIN_X) % BIN_Y, z = z0 + at / (BIN_X * BIN_Y);
        if (x >= n || y >= n || z >= n) continue;
        const uint32_t e = box[at], et = box[BIN_VOXELS + at], mx = box[2 * BIN_VOXELS + at];
        const uint32_t my = box[3 * BIN_VOXELS + at], mz = box[4 * BIN_VOXELS + at], mass = box[5 * BIN_VOXELS + at];
        if ((e | et | mx | my | mz | mass) == 0u) continue;
        const uint32_t index = (z * n + y) * n + x;
        uint32_t* em = a.emission + 2u * index;
        uint32_t* mo = reinterpret_cast<uint32_t*>(a.momentum) + 3u * index;
        em[0] += e;
        em[1] += et;
        mo[0] += mx;
        mo[1] += my;
        mo[2] += mz;
        a.mass[index] += mass;
    }
}

SplatBins splat_bin_layout(uint32_t size) {
    SplatBins g{};
    g.nbx = (size + BIN_X - 1u) / BIN_X;
    g.nby = (size + BIN_Y - 1u) / BIN_Y;
    const uint64_t bricks = (uint64_t)g.nbx * g.nby * ((size + BIN_Z - 1u) / BIN_Z);
    g.bricks = bricks <= kSplatBinMaxBricks ? (uint32_t)bricks : 0u;
    return g;
}

uint64_t splat_bin_table_words(uint32_t size) {
    const SplatBins g = splat_bin_layout(size);
    return g.bricks ? (uint64_t)kSplatBinGroups * g.bricks + 2u * g.bricks + 1u : 0u;
}

hipError_t launch_splat_binned(const SplatArgs& a, SplatBins g, hipStream_t stream) {
    const uint32_t count = a.p.count;
    if (g.bricks == 0 || count == 0) return hipErrorInvalidValue;
    g.groups = std::max(1u, std::min(kSplatBinGroups, (count + 4095u) / 4096u));
    g.chunk = (uint32_t)(((uint64_t)count + g.groups - 1u) / g.groups);
    splat_bin_count_kernel<<<dim3(g.groups), BIN_THREADS, 0, stream>>>(a, g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    splat_bin_scan_kernel<<<dim3((g.bricks + SCAN_BRICKS - 1u) / SCAN_BRICKS), BIN_THREADS, 0, stream>>>(g);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    splat_bin_scatter_kernel<<<dim3(g.groups), BIN_THREADS, 0, stream>>>(a, g);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    splat_bin_brick_kernel<<<dim3(g.bricks), BIN_BRICK_THREADS, 0, stream>>>(a, g);
    return hipGetLastError();
}

namespace {

// Onesweep at every size above one block: rocprim's default picks a merge
// sort up to 2^20 items, several times slower for these keys.
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;

int key_bits(uint32_t size) {
    const uint64_t m = (uint64_t)size + 1u;
    int bits = 1;
    while ((1ull << bits) <= m * m * m) ++bits;
    return bits;
}

}  // namespace

// Storage for the largest key (32 bits), so one allocation serves every
// volume size.
size_t splat_sort_temp_bytes(uint32_t count) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs<SortConfig>(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                (const uint32_t*)nullptr, (uint32_t*)nullptr, count, 0, 32);
    return bytes;
}

const uint32_t kSplatBricks[kSplatBrickShapes][3] = {{16, 8, 4}, {8, 8, 4}, {8, 8, 8},   {16, 16, 4},
                                                     {32, 8, 4}, {16, 8, 8}, {32, 16, 4}, {16, 16, 8}};

bool splat_brick_supported(const uint32_t shape[3]) {
    for (const auto& b : kSplatBricks)
        if (b[0] == shape[0] && b[1] == shape[1] && b[2] == shape[2]) return true;
    return false;
}

uint64_t splat_gather_bricks(uint32_t size, const uint32_t shape[3]) {
    return (uint64_t)((size + shape[0] - 1) / shape[0]) * ((size + shape[1] - 1) / shape[1]) *
           ((size + shape[2] - 1) / shape[2]);
}

uint64_t splat_gather_bricks_max(uint32_t size) {
    uint64_t most = 0;
    for (const auto& b : kSplatBricks) most = std::max(most, splat_gather_bricks(size, b));
    return most;
}

namespace {

template <uint32_t X, uint32_t Y, uint32_t Z>
hipError_t launch_bricks(const SplatArgs& a, const SplatSortBuffers& b, const SplatGather& g, hipStream_t stream) {
    const uint32_t shape[3] = {X, Y, Z};
    splat_brick_kernel<X, Y, Z><<<dim3((uint32_t)splat_gather_bricks(a.size, shape)), BRICK_THREADS, 0, stream>>>(a, b, g);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_splat_sorted(const SplatArgs& a, const SplatSortBuffers& b, const SplatGather& g,
                               hipStream_t stream) {
    const uint32_t count = a.p.count;
    const uint32_t blocks = (count + 255u) / 256u;
    splat_key_kernel<<<dim3(blocks), 256, 0, stream>>>(a, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t temp = b.temp_bytes;
    e = rocprim::radix_sort_pairs<SortConfig>(b.temp, temp, b.keys_in, b.keys_out, b.order_in, b.order_out, count, 0,
                                              key_bits(a.size), stream);
    if (e != hipSuccess) return e;
    splat_stage_kernel<<<dim3(blocks), 256, 0, stream>>>(a, b, g);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t x = g.brick[0], y = g.brick[1], z = g.brick[2];
    if (x == 16 && y == 8 && z == 4) return launch_bricks<16, 8, 4>(a, b, g, stream);
    if (x == 8 && y == 8 && z == 4) return launch_bricks<8, 8, 4>(a, b, g, stream);
    if (x == 8 && y == 8 && z == 8) return launch_bricks<8, 8, 8>(a, b, g, stream);
    if (x == 16 && y == 16 && z == 4) return launch_bricks<16, 16, 4>(a, b, g, stream);
    if (x == 32 && y == 8 && z == 4) return launch_bricks<32, 8, 4>(a, b, g, stream);
    if (x == 16 && y == 8 && z == 8) return launch_bricks<16, 8, 8>(a, b, g, stream);
    if (x == 32 && y == 16 && z == 4) return launch_bricks<32, 16, 4>(a, b, g, stream);
    if (x == 16 && y == 16 && z == 8) return launch_bricks<16, 16, 8>(a, b, g, stream);
    return hipErrorInvalidValue;
}

}  // namespace garg
