// Kernel argument blocks and the launchers each .hip file provides to the
// C-ABI layer (api.cpp). Arguments travel by value in the kernarg segment.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gargantua.h"
#include "device_common.hpp"

namespace garg {

struct MarchImageArgs {
    garg_march_uniforms uniforms;
    Scene scene;
    h4* color;
    h4* debug;
    float* depth;
    garg_march_counters* counters;
    uint32_t extent_w, extent_h;
};

struct ProbeArgs {
    garg_march_uniforms uniforms;
    Scene scene;
    const garg_ray_probe* probes;
    garg_ray_probe_result* results;
    uint32_t count;
};

hipError_t launch_march_image(const MarchImageArgs& a, uint32_t integrator, hipStream_t stream);
// Adds the counters at `src` to those at `a` and at `b` (one thread).
hipError_t launch_add_counters(garg_march_counters* a, garg_march_counters* b, const garg_march_counters* src,
                               hipStream_t stream);
// Workgroups (one wave each) of the one-launch-per-frame march that one
// compute unit holds at once, for `integrator`; 0 if unknown.
uint32_t march_image_resident_blocks(uint32_t integrator);
hipError_t launch_probe(const ProbeArgs& a, uint32_t integrator, hipStream_t stream);

// One request of the persistent march as the device sees it.
struct BatchRequest {
    garg_march_uniforms uniforms;
    Scene scene;
    h4* color;
    h4* debug;
    float* depth;
    garg_march_counters* counters;
    uint32_t tiles_x;     // 8 x 8 tiles per row of the image
    uint32_t ray_begin;   // first ray of this request, tile order
    uint64_t queue_begin; // offset of this request in the batch's ray queue
    uint32_t ray_count;   // rays of this request in the batch; its queue
                          // range is padded to a multiple of 64
    uint32_t pad;
};

// Requests of one persistent launch; the LDS counter table bounds their number.
#define GARG_BATCH_MAX_REQUESTS 256

struct BatchArgs {
    const BatchRequest* requests;   // sorted by queue_begin, first at 0
    uint32_t request_count;
    uint32_t kind;                  // BATCH_SCHWARZSCHILD (every request has spin 0), BATCH_KERR,
                                    // BATCH_KERR_SCHILD or BATCH_KERR_POLAR (every request has spin > 0)
    uint64_t total_rays;
    unsigned long long* queue;      // refill policy: eight {head, end} pairs, 128
                                    // bytes apart; tile policy: one head
    garg_march_counters* totals;
    uint32_t compaction_period;
    uint32_t refill_threshold;
    uint32_t aging;                 // whole-tile claims: raise a tile's issue priority as it ages
                                    // (always for Kerr; for Schwarzschild when a budget exceeds
                                    // the interactive 256 steps)
};

// The batched march's ray state machines.
enum { BATCH_SCHWARZSCHILD = 0, BATCH_KERR = 1, BATCH_KERR_SCHILD = 2, BATCH_KERR_POLAR = 3 };

hipError_t launch_march_batch(const BatchArgs& a, uint32_t grid, uint32_t block, hipStream_t stream);
uint32_t march_batch_resident_blocks(uint32_t kind, bool aging, bool tiles, uint32_t block);

// Particles, structure of arrays.
struct ParticleView {
    float* px;
    float* py;
    float* pz;
    float* vx;
    float* vy;
    float* vz;
    float* temperature;
    uint32_t* respawns;
    uint32_t* dispatches;  // one word: sim dispatches since seeding
    uint32_t count;
};

struct SimArgs {
    ParticleView p;
    uint32_t seed_lo, seed_hi;
    uint32_t potential;
    uint32_t generator;
    float sigma_z;
    float dt, alpha, temperature_scale;
    const float* inverse_cdf;
    uint32_t inverse_cdf_count;
};

hipError_t launch_seed(const SimArgs& a, hipStream_t stream);
// `dispatches` consecutive sim dispatches in one launch.
hipError_t launch_step(const SimArgs& a, uint32_t dispatches, hipStream_t stream);
hipError_t launch_pack_aos(const ParticleView& p, garg_particle* out, hipStream_t stream);
hipError_t launch_unpack_aos(const ParticleView& p, const garg_particle* in, hipStream_t stream);

struct SplatArgs {
    ParticleView p;
    uint32_t* emission;   // 2 per voxel
    int32_t* momentum;    // 3 per voxel
    uint32_t* mass;       // 1 per voxel
    uint32_t size;
    float half_xy, half_z;
    float peak_temperature, emission_scale;
    uint32_t exact, frame;
    uint32_t seed_lo, seed_hi;
    const uint32_t* frame_counter;  // non-null: the frame is this word (read at run time)
};

struct ResolveArgs {
    uint32_t* emission;
    int32_t* momentum;
    uint32_t* mass;
    Voxel* voxels;
    const float4* blackbody;
    uint32_t size;
    float half_xy, half_z;
    float peak_temperature;
    uint32_t clear;
};

hipError_t launch_splat_bin(const SplatArgs& a, hipStream_t stream);

// Buffers of the sorted exact deposit: keys and particle order (in and out
// of the radix sort, `count` words each) and the sort's temporary storage.
// The exact deposit's per-particle buffers (in the particle store): cell
// keys and particle indices before and after the sort, the sort's storage,
// and each particle's 32-byte record (fractions x y z, energy; temperature
// ratio, velocity x y z) in particle order and in sorted order.
struct SplatSortBuffers {
    uint32_t* keys_in;
    uint32_t* keys_out;
    uint32_t* order_in;
    uint32_t* order_out;
    void* temp;
    size_t temp_bytes;
    float4* records;
    float4* sorted;
};
// The exact deposit's per-volume tables: the first sorted index and one past
// the last of each of the (size + 1)^3 cells (trusted only where the sorted
// keys confirm them), and one word per voxel brick holding the stamp of the
// last deposit that touched it.
struct SplatGather {
    uint32_t* cells;
    uint32_t* ends;
    uint32_t* bricks;
    uint32_t stamp;
    uint32_t brick[3];  // voxels per brick on x, y, z: one of kSplatBricks
};
size_t splat_sort_temp_bytes(uint32_t count);
// The brick shapes the exact deposit is built for; the first is the default.
constexpr int kSplatBrickShapes = 8;
extern const uint32_t kSplatBricks[kSplatBrickShapes][3];
bool splat_brick_supported(const uint32_t shape[3]);
uint64_t splat_gather_bricks(uint32_t size, const uint32_t shape[3]);
// Brick stamps a volume of this size needs for its finest supported shape.
uint64_t splat_gather_bricks_max(uint32_t size);
hipError_t launch_splat_sorted(const SplatArgs& a, const SplatSortBuffers& b, const SplatGather& g,
                               hipStream_t stream);
// The one-corner deposit by a counting sort over 16 x 16 x 8 voxel bricks
// (splat_sorted.hip): per-particle keys and two 16-byte records (the
// particle store's sort buffers), and per-volume tables of
// kSplatBinGroups x bricks counts, the bricks' totals and their bricks + 1
// starts. Volumes of more than kSplatBinMaxBricks bricks (sides past 320
// voxels) deposit directly.
constexpr uint32_t kSplatBinBrick[3] = {16, 16, 8};
constexpr uint32_t kSplatBinMaxBricks = 16384;
constexpr uint32_t kSplatBinGroups = 256;
struct SplatBins {
    uint32_t* keys;
    uint4* records;
    uint32_t* counts;
    uint32_t* totals;
    uint32_t* starts;
    uint32_t bricks, nbx, nby;  // bricks = 0: the volume is too large for this path
    uint32_t groups, chunk;     // set at launch from the particle count
};
// The brick grid of a volume side (bricks = 0 past kSplatBinMaxBricks), and
// the words its tables need (0 then).
SplatBins splat_bin_layout(uint32_t size);
uint64_t splat_bin_table_words(uint32_t size);
hipError_t launch_splat_binned(const SplatArgs& a, SplatBins bins, hipStream_t stream);
hipError_t launch_resolve(const ResolveArgs& a, hipStream_t stream);
hipError_t launch_fill_zero(void* ptr, uint64_t bytes, hipStream_t stream);
hipError_t launch_atomic_rate(uint32_t* buffer, uint32_t words, uint32_t adds, uint32_t threads, hipStream_t stream);
// `threads` x `iterations` x 32 independent FP32 fused multiply-adds in
// operand form `form` (0: two VGPRs and a literal, 1: three VGPRs, 2: SGPRs).
hipError_t launch_valu_rate(float* out, uint32_t form, uint32_t threads, uint32_t iterations, hipStream_t stream);

// FP64 sweep tables: table[i] for the sweep from `radius` over b at table
// coordinate i / (n - 1) between low and high; kind 0 = sweep to infinity,
// kind 1 = full escape sweep (turn plus outgoing leg).
hipError_t launch_sweep_table(float* out, uint32_t n, double radius, double low, double high, uint32_t intervals,
                              int kind, hipStream_t stream);
hipError_t launch_far_field(double* b, double* alpha, double* sweep, uint32_t n, double r_obs, double b_lo, double b_hi,
                            uint32_t intervals, hipStream_t stream);
hipError_t launch_shadow(double* edge, double* xi, uint32_t azimuths, double spin, double r_obs, double inclination,
                         uint32_t rounds, uint32_t rays_per_round, uint32_t step_cap, hipStream_t stream);

struct TonemapArgs {
    const h4* color;
    const h4* debug;
    uint8_t* out;
    uint32_t pixels;
    float exposure;
    uint32_t debug_view;
    float drift_budget;
};
hipError_t launch_tonemap(const TonemapArgs& a, hipStream_t stream);

struct BakeArgs {
    garg_march_uniforms uniforms;
    Scene scene;
    h4* samples;         // capacity x rays, sample k of ray i at k * rays + i
    uint4* headers;      // count, falls in, packed direction xy, packed z + depth
    uint32_t* overflow;
    uint32_t row0, rows;
    uint32_t capacity;
};

struct WalkArgs {
    garg_march_uniforms uniforms;
    Scene scene;
    const h4* samples;
    const uint4* headers;
    h4* color;
    float* depth;
    uint32_t capacity;
};

// The FP64 probe's policy: garg_geodesic_settings widened to double.
struct ProbeSettings64 {
    double step_factor, step_min, step_max, capture_radius, escape_radius;
    uint64_t step_cap;
    double empty_step_factor, empty_step_max;
};
hipError_t launch_probe_fp64(const garg_ray_probe* probes, uint32_t count, const ProbeSettings64& g,
                             garg_ray_probe_fp64_result* out, hipStream_t stream);

// The FP64 lensed render of a resolved volume: one pixel per lane, linear
// rgb (3 doubles per pixel, row major) into `rgb`.
struct Render64 {
    const Voxel* voxels;
    uint32_t size;
    double half_xy, half_z;
    double position[3], right[3], up[3], forward[3];
    double tan_half_fov[2];
    uint32_t width, height;
    double opacity;
    uint32_t redshift;
    ProbeSettings64 settings;
    double* rgb;
};
hipError_t launch_render_fp64(const Render64& a, hipStream_t stream);

hipError_t launch_bake(const BakeArgs& a, hipStream_t stream);
hipError_t launch_walk(const WalkArgs& a, hipStream_t stream);

}  // namespace garg
