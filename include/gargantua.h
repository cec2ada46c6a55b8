/*
 * gargantua for MI355X: the C ABI of the per-frame compute passes.
 *
 * One shared library, libgargantua_hip.so, built for gfx950. Every entry
 * point takes plain pointers and sizes, returns a status (0 = ok) and leaves
 * a one-sentence description of any failure in garg_last_error(). Nothing
 * here exits the process: the caller maps failures to its own exit codes.
 *
 * The boundary replaces the reference's Metal pass objects and the shared C
 * header they bind (Sources/ShaderTypes/include/ShaderTypes.h). Each entry
 * point names the reference interface it stands in for. Launch entry points
 * take a stream (NULL = the context's default stream) the way the
 * reference's encode calls take an MTLCommandBuffer; they enqueue and
 * return, and garg_stream_sync waits, like waitUntilCompleted.
 *
 * Threading: a context and everything created from it may be used by one
 * host thread at a time.
 */
#ifndef GARGANTUA_H
#define GARGANTUA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status, enums ------------------------------------------------------ */

enum garg_status {
    GARG_OK = 0,
    GARG_E_ARGUMENT = 1,  /* invalid argument; nothing was enqueued */
    GARG_E_DEVICE = 2,    /* the HIP runtime reported an error */
    GARG_E_BUDGET = 3,    /* refused: the request does not fit the memory budget */
    GARG_E_STATE = 4      /* the object is not in a state that allows the call */
};

/* Ray outcomes and potentials (ShaderTypes.h:22-25,72-76). */
enum { GARG_RAY_CAPTURED = 0, GARG_RAY_ESCAPED = 1, GARG_RAY_EXHAUSTED = 2 };
enum { GARG_POTENTIAL_NEWTONIAN = 0, GARG_POTENTIAL_PACZYNSKI_WIITA = 1 };

/* Particle generators: the reference seeding (event 0 from the steady
 * profile's inverse CDF, z = 0.05 r N, sim.metal:6-34) and the benchmark
 * scene generator (area-uniform radius in the feeding annulus, z = sigma N). */
enum { GARG_GENERATOR_DISK = 0, GARG_GENERATOR_ANNULUS = 1 };

/* Integrator forms of the march. REFERENCE evaluates the reference's
 * formulas as written (Geodesic.h, Kerr.h); FAST is this library's form
 * (Kerr: one sincos per step, angle addition for the inner stages, three
 * reciprocals per stage, drift gauges shared with the next stage;
 * Schwarzschild: one reciprocal of r per stage in place of four divisions),
 * with TwoSum-compensated state accumulation when the uniforms ask for it.
 * KERR_SCHILD integrates Kerr rays in Cartesian Kerr-Schild coordinates,
 * whose metric is regular on the spin axis: the form the reference names as
 * future work for its polar seam (docs/lab/07-kerr.md:33-45). It takes the
 * same launch, step policy (without the pole factor), gas and sky as the
 * Boyer-Lindquist forms; at spin 0 it is the FAST Schwarzschild form.
 * POLAR marches a Kerr ray in the FAST form until sin(theta) first falls
 * below 0.25, or the ray is inside r+ + 0.5 falling faster than
 * dr/dlambda = -2, and in the KERR_SCHILD form from there to its end,
 * converting the state once, in double precision: regular on the axis and
 * across the horizon at about the cost of the FAST form, since only rays
 * that pass near the axis or plunge leave it; at spin 0 it is the FAST
 * Schwarzschild form.
 * The batched march runs the FAST form, or for spinning requests the
 * KERR_SCHILD or POLAR form when garg_batch_config.kerr_schild asks for it. */
enum {
    GARG_INTEGRATOR_REFERENCE = 0,
    GARG_INTEGRATOR_FAST = 1,
    GARG_INTEGRATOR_KERR_SCHILD = 2,
    GARG_INTEGRATOR_POLAR = 3
};

typedef struct garg_ctx garg_ctx;
typedef struct garg_stream garg_stream;
typedef struct garg_event garg_event;
typedef struct garg_particles garg_particles;
typedef struct garg_volume garg_volume;
typedef struct garg_tables garg_tables;
typedef struct garg_targets garg_targets;
typedef struct garg_bake garg_bake;
typedef struct garg_graph garg_graph;

/* ---- shared records (ShaderTypes.h) -------------------------------------- */

/* One particle, 32 bytes: the export layout of the particle store
 * (ShaderTypes.h:15-20). The device store itself is structure of arrays. */
typedef struct {
    float position[3];
    float velocity[3];
    float temperature;
    uint32_t respawns;
} garg_particle;

/* Geodesic integration policy (ShaderTypes.h:60-70). */
typedef struct {
    float step_factor;
    float step_min;
    float step_max;
    float capture_radius;
    float escape_radius;
    uint32_t step_cap;
    float empty_step_factor;
    float empty_step_max;
} garg_geodesic_settings;

/* Per-render uniforms (ShaderTypes.h:108-144), plus two fields of this
 * library: the z half extent of a non-cubic volume and the compensated
 * accumulation switch. */
typedef struct {
    float camera_position[3];
    float camera_right[3];
    float camera_up[3];
    float camera_forward[3];
    float tan_half_fov[2];      /* tan(fov / 2), scaled by aspect in x */
    uint32_t resolution[2];     /* full image size in pixels */
    uint32_t tile_origin[2];    /* pixel offset of the tile being rendered */
    float volume_half_extent;
    float opacity_scale;
    float star_brightness;
    float drift_budget;
    uint32_t star_seed;
    uint32_t redshift;
    float bake_spacing;
    uint32_t bake_capacity;
    float sphere_radius;        /* 0 integrates every ray from the camera */
    float camera_radius;
    float sphere_max_b;
    float sphere_table_max_b;
    float camera_max_b;
    float sky_min_b;
    float sky_max_b;
    float spin;                 /* 0 renders Schwarzschild */
    float jitter[2];
    float volume_half_extent_z; /* 0 means equal to volume_half_extent */
    uint32_t compensated;       /* 1: TwoSum-compensated state accumulation */
    garg_geodesic_settings geodesic;
} garg_march_uniforms;

/* Validation probes (ShaderTypes.h:79-92). */
typedef struct {
    float origin[3];
    float direction[3];
} garg_ray_probe;

typedef struct {
    float position[3];
    float direction[3];
    float drift;
    uint32_t outcome;
    uint32_t steps;
    float energy;
    float angular_momentum;
} garg_ray_probe_result;

/* Per-launch ray statistics (ShaderTypes.h:157-162), plus the sum of
 * per-ray RK4 steps, the unit of the geodesic steps/s metric. */
typedef struct {
    uint32_t rays;
    uint32_t over_drift_budget;
    uint32_t exhausted;
    uint32_t captured;
    uint64_t steps;
} garg_march_counters;

typedef struct {
    char name[64];
    char arch[32];
    uint32_t compute_units;
    uint32_t clock_khz;
    uint64_t total_bytes;
    uint64_t free_bytes;
    uint64_t budget_bytes;  /* what this context will allocate at most */
} garg_device_info;

/* ---- context, streams, events (Gpu/GpuContext.swift:8-95) ---------------- */

const char* garg_last_error(void);
const char* garg_version(void);
/* budget_bytes = 0 takes 90% of free device memory. */
int garg_ctx_create(int device, uint64_t budget_bytes, garg_ctx** out);
int garg_ctx_destroy(garg_ctx* ctx);
int garg_device_info_get(garg_ctx* ctx, garg_device_info* out);
/* Bytes this context holds now. */
int garg_ctx_allocated(garg_ctx* ctx, uint64_t* bytes);

int garg_stream_create(garg_ctx* ctx, garg_stream** out);
int garg_stream_destroy(garg_stream* stream);
int garg_stream_sync(garg_ctx* ctx, garg_stream* stream);
int garg_event_create(garg_ctx* ctx, garg_event** out);
int garg_event_destroy(garg_event* event);
int garg_event_record(garg_event* event, garg_stream* stream);
int garg_event_sync(garg_event* event);
int garg_event_elapsed_ms(garg_event* start, garg_event* end, float* ms);

/* ---- particles: ParticleSystem.swift:8-70, SimPass.swift:15-22 ----------- */

int garg_particles_create(garg_ctx* ctx, uint32_t count, uint64_t seed, uint32_t potential, garg_particles** out);
int garg_particles_destroy(garg_particles* particles);
/* seedParticles (sim.metal:36-42); sigma_z is used by the annulus generator. */
int garg_particles_seed(garg_particles* particles, garg_stream* stream, uint32_t generator, float sigma_z);
/* stepParticles (sim.metal:44-79), `dispatches` times; each is SIM_SUBSTEPS
 * substeps. Consecutive dispatches run in one launch (up to 256 per launch)
 * with each particle held in registers: the same bits as one launch each. */
int garg_particles_step(garg_particles* particles, garg_stream* stream, uint32_t dispatches);
int garg_particles_reseed(garg_particles* particles, uint64_t seed);
/* Synchronous AoS export / import of `count` records (Validate.swift:88-100). */
int garg_particles_read(garg_particles* particles, garg_particle* host, uint32_t count);
int garg_particles_write(garg_particles* particles, const garg_particle* host, uint32_t count);
/* 64-bit FNV-1a over the 32-byte AoS records (Validate.swift:378-387). */
int garg_particles_checksum(garg_particles* particles, uint64_t* fnv1a);
/* Sim dispatches since the last seed (each garg_particles_step dispatch adds
 * one, on the device). */
int garg_particles_dispatches(garg_particles* particles, uint32_t* out);

/* ---- volume: Gpu/Volume.swift:14-43, SplatPass.swift:34-76 -------------- */

/* half_extent_z = 0 makes the cube of half extent `half_extent`. */
int garg_volume_create(garg_ctx* ctx, uint32_t size, float half_extent, float half_extent_z, garg_volume** out);
int garg_volume_destroy(garg_volume* volume);
/* Volume.encodeClear: zero the fixed point sums. */
int garg_volume_clear(garg_volume* volume, garg_stream* stream);
/* splatParticles (splat.metal:44-88): exact = 1 deposits into all eight
 * corners, 0 picks one corner per particle from pcg4d(i, frame, seed).
 * frame = GARG_FRAME_AUTO takes the frame from the particle store's dispatch
 * count when the kernel runs, so a captured graph of step + splat + resolve
 * draws a new frame on every replay (the reference passes a per-call
 * counter, SplatPass.swift:18). The exact deposit uses the particle store's
 * sort buffers and the volume's cell tables, so two exact deposits that
 * share either must not run concurrently; inside a graph capture it takes
 * the direct atomic path, which needs neither. */
#define GARG_FRAME_AUTO 0xffffffffu
int garg_splat_bin(garg_volume* volume, garg_particles* particles, garg_stream* stream, uint32_t exact, uint32_t frame);
/* No counterpart in the reference (its deposit is one atomic pass): the
 * brick shape of the exact deposit's gather, for the autotuner. Each brick of
 * x * y * z voxels is one 512-thread workgroup that sums its corners in LDS
 * (24 bytes per voxel) and owns those voxels' sums, so the shape trades LDS
 * per workgroup, the halo of cells read around a brick and the number of
 * workgroups. Shapes: 16x8x4 (the default), 8x8x4, 8x8x8, 16x16x4, 32x8x4,
 * 16x8x8, 32x16x4, 16x16x8; GARG_E_ARGUMENT for any other. Every shape gives
 * the same sums, bit for bit (integers added modulo 2^32). */
int garg_volume_set_splat_brick(garg_volume* volume, uint32_t x, uint32_t y, uint32_t z);
int garg_volume_splat_brick(garg_volume* volume, uint32_t shape[3]);
/* No counterpart in the reference (its deposit is one pass of device-scope
 * atomics, splat.metal:25-88): how garg_splat_bin deposits into this volume.
 * GARG_SPLAT_SORTED (the default) sorts instead of adding to random words:
 * the exact deposit by cell with the library radix sort (outside graph
 * captures), the one-corner deposit by 16x16x8 voxel brick with a counting
 * sort of the library's own kernels (captured or not; volume sides up to
 * 320, beyond which it deposits directly). GARG_SPLAT_ATOMIC always runs the
 * reference's atomic pass. Both give the same sums, bit for bit. The
 * one-corner sort uses the particle store's sort buffers, like the exact one:
 * deposits that share a particle store must not run concurrently. */
#define GARG_SPLAT_SORTED 0u
#define GARG_SPLAT_ATOMIC 1u
int garg_volume_set_splat_method(garg_volume* volume, uint32_t method);
int garg_volume_splat_method(garg_volume* volume, uint32_t* method);
/* resolveVolume (splat.metal:93-120); clear_sums = 1 also zeroes the sums. */
int garg_splat_resolve(garg_volume* volume, garg_stream* stream, uint32_t clear_sums);
/* Synchronous reads: sums (2 u32, 3 i32, 1 u32 per voxel, z-major) and the
 * resolved fields (rgba16f emission + density, rgba16f velocity). */
int garg_volume_read_sums(garg_volume* volume, uint32_t* emission, int32_t* momentum, uint32_t* mass);
int garg_volume_read_fields(garg_volume* volume, uint16_t* emission_rgba16f, uint16_t* velocity_rgba16f);
int garg_volume_write_fields(garg_volume* volume, const uint16_t* emission_rgba16f, const uint16_t* velocity_rgba16f);

/* ---- sweep tables: Gpu/DeflectionTables.swift:10-82 ---------------------- */

int garg_tables_create(garg_ctx* ctx, garg_stream* stream, float sphere_radius, garg_tables** out);
int garg_tables_destroy(garg_tables* tables);
/* DeflectionTables.update: regenerate the camera and sky tables in FP64 on
 * the GPU when the camera radius moves. */
int garg_tables_update(garg_tables* tables, garg_stream* stream, float camera_radius);
/* DeflectionTables.apply: enable the jump in a set of uniforms. */
int garg_tables_apply(garg_tables* tables, garg_march_uniforms* uniforms);
int garg_tables_read(garg_tables* tables, float* sphere, float* camera, float* sky);

/* ---- targets: MarchPass.makeTargets (MarchPass.swift:42-50) --------------- */

int garg_targets_create(garg_ctx* ctx, uint32_t width, uint32_t height, garg_targets** out);
int garg_targets_destroy(garg_targets* targets);
/* Synchronous reads; NULL pointers are skipped. colour and debug are
 * rgba16f per pixel (debug = drift, steps, outcome, 1), depth is float. */
int garg_targets_read(garg_targets* targets, uint16_t* color_rgba16f, uint16_t* debug_rgba16f, float* depth);
int garg_targets_read_counters(garg_targets* targets, garg_march_counters* out);
int garg_targets_reset_counters(garg_targets* targets, garg_stream* stream);

/* ---- march: MarchPass.encodeImage / encodeProbes (MarchPass.swift:103-168) */

/* marchImage (march.metal:194-288): one thread per pixel, 8x8 groups, one
 * launch over the rectangle [tile_origin, tile_origin + extent) of a
 * targets image of size `resolution`. extent 0 means the whole image. The
 * counters of `targets` accumulate. Spinning scenes keep the escape radius
 * and the camera within r = KERR_MAX_RADIUS (GARG_E_ARGUMENT beyond), as do
 * spinning probes their origins and every request of garg_march_batch. */
int garg_march_image(garg_ctx* ctx, garg_stream* stream, const garg_march_uniforms* uniforms, garg_volume* volume,
                     garg_tables* tables, garg_targets* targets, uint32_t extent_w, uint32_t extent_h,
                     uint32_t integrator);

/* probeRays (march.metal:290-341), synchronous: host probes in, host results out. */
int garg_probe(garg_ctx* ctx, garg_stream* stream, const garg_ray_probe* probes, uint32_t count,
               const garg_march_uniforms* uniforms, garg_tables* tables, uint32_t integrator,
               garg_ray_probe_result* results);

/* No counterpart in the reference. garg_probe for callers that cannot
 * allocate device memory (tight validation or serving loops, harnesses that
 * forbid hipMalloc): the same checks, kernel and results, but the device
 * staging of the probes and results is a fixed array inside this library
 * (1.1 MB per device, held until the process ends) instead of a context
 * allocation, and the call allocates nothing. At most GARG_PROBE_NOALLOC_MAX
 * probes per call (GARG_E_ARGUMENT beyond). Every call on one device shares
 * that device's array: two calls that overlap on one device (from two host
 * threads, or two contexts) race and corrupt each other's results. */
#define GARG_PROBE_NOALLOC_MAX 16384u
int garg_probe_noalloc(garg_ctx* ctx, garg_stream* stream, const garg_ray_probe* probes, uint32_t count,
                       const garg_march_uniforms* uniforms, garg_tables* tables, uint32_t integrator,
                       garg_ray_probe_result* results);

/* ---- batched persistent march (no reference counterpart; nearest is the
 * reference's one launch per frame, MarchPass.encodeFrame) ------------------ */

typedef struct {
    garg_march_uniforms uniforms;
    garg_targets* targets;  /* full-resolution targets of this request */
    garg_tables* tables;    /* NULL: no sphere jump (always NULL for spin > 0) */
    garg_volume* volume;    /* NULL: no gas */
    uint32_t ray_begin;     /* first ray of the request in tile order */
    uint32_t ray_count;     /* rays to march from ray_begin; 0 = to the end */
} garg_march_request;

typedef struct {
    uint32_t block;             /* threads per workgroup: 64, 128 or 256 */
    uint32_t waves_per_simd;    /* waves per SIMD of the persistent grid (capped at what
                                   fits at once); 0 = 4 */
    uint32_t compaction_period; /* RK4 steps between refill checks */
    uint32_t refill_threshold;  /* refill a wave when at most this many of 64 lanes are
                                   live; 0 = whole-tile claims: a wave marches 64 rays to
                                   the end before it takes the next 64 */
    uint32_t kerr_schild;       /* spinning requests: 0 = the Boyer-Lindquist FAST form,
                                   1 = the Kerr-Schild form (regular on the spin axis),
                                   2 = the POLAR form (Kerr-Schild near the axis and on plunges only) */
} garg_batch_config;

/* Rays of a request in tile order: 8x8 pixel tiles row-major, 64 rays each. */
uint32_t garg_request_rays(uint32_t width, uint32_t height);
int garg_batch_config_default(garg_ctx* ctx, garg_batch_config* out);
/* Persistent launches marching every requested ray of every request with
 * work stealing and either live-ray refill or whole-tile claims (see
 * garg_batch_config): one launch for the spinning requests
 * and one for the Schwarzschild ones, at most 256 requests. Each request is
 * queued as up to four bands of whole tiles, middle bands first and
 * interleaved across requests; no result depends on the queue order. With
 * whole-tile claims, a lone request of a whole image at most two and a half
 * rounds of waves deep takes the one-launch-per-frame kernel instead, whose
 * workgroups fit more waves per SIMD (same bits). Each request's targets
 * counters accumulate; `total` (optional) receives the sums over the call,
 * read back synchronously. */
int garg_march_batch(garg_ctx* ctx, garg_stream* stream, const garg_march_request* requests, uint32_t count,
                     const garg_batch_config* config, garg_march_counters* total);

/* ---- bake / walk: BakePass.swift:21-102, bake.metal ----------------------- */

int garg_bake_create(garg_ctx* ctx, uint32_t width, uint32_t height, uint32_t capacity, garg_bake** out);
int garg_bake_destroy(garg_bake* bake);
/* Bytes the bake buffers of this size occupy (samples + headers). */
uint64_t garg_bake_bytes(uint32_t width, uint32_t height, uint32_t capacity);
/* bakeRays over rows [row0, row0 + rows) (encodeBakeChunk); row0 = 0 resets the overflow count. */
int garg_bake_rows(garg_bake* bake, garg_stream* stream, const garg_march_uniforms* uniforms, garg_tables* tables,
                   uint32_t row0, uint32_t rows);
/* walkRays (bake.metal:94-144) into `targets`. */
int garg_bake_walk(garg_bake* bake, garg_stream* stream, const garg_march_uniforms* uniforms, garg_volume* volume,
                   garg_targets* targets);
int garg_bake_overflow(garg_bake* bake, uint32_t* rays);
/* Synchronous reads of the per-ray headers (4 u32 each) and of the samples
 * (rgba16f, sample k of ray i at k * rays + i). */
int garg_bake_read(garg_bake* bake, uint32_t* headers, uint16_t* samples);

/* ---- present: the PBR Neutral curve (present.metal:19-58) ---------------- */

/* Exposure, tone curve and sRGB encoding into 8-bit RGBA, synchronous.
 * debug_view = 1 draws the drift heat map instead. */
int garg_tonemap(garg_targets* targets, garg_stream* stream, float exposure, uint32_t debug_view, float drift_budget,
                 uint8_t* rgba8);

/* ---- FP64 GPU kernels ------------------------------------------------------ */

/* Far-field deflection table in FP64: for b_k log-spaced on [b_lo, b_hi],
 * alpha_k = total deflection of a ray from infinity and sweep_k = sweep from
 * r_obs to infinity, midpoint rule after u = u_max sin^2 (Deflection.swift:23-39). */
int garg_deflection_table(garg_ctx* ctx, garg_stream* stream, double r_obs, uint32_t count, double b_lo, double b_hi,
                          uint32_t intervals, double* b_out, double* alpha_out, double* sweep_out);

/* Kerr shadow boundary in FP64 along `azimuths` screen directions: for each
 * screen azimuth psi, the angular radius of the capture boundary seen from a
 * static observer at (r_obs, inclination), found by multi-section over
 * `rounds` rounds of `rays_per_round` rays (Validate.swift:125-150 for the
 * multi-section, KerrGeodesic.swift:249-262 for the launch convention).
 * edge_out[k] = boundary aim angle, xi_out[k] = L / E there (optional). */
int garg_shadow_boundary(garg_ctx* ctx, garg_stream* stream, double spin, double r_obs, double inclination_deg,
                         uint32_t azimuths, uint32_t rounds, uint32_t rays_per_round, uint32_t step_cap,
                         double* edge_out, double* xi_out);

/* The reference's FP64 Schwarzschild integrator (Oracle/Schwarzschild.swift:
 * 76-97 launch, 121-136 RK4, 174-184 step length with the empty-leg coarse
 * step, 200-220 the loop and its energy / angular momentum drift) on the GPU
 * in double precision, one ray per lane, for the probes' launches and the
 * settings widened to double. The reference judges its parity and sky
 * goldens against that integrator on the CPU (Validate.swift:179-200,
 * 293-320); `gargantua validate` judges them against this. Synchronous; at
 * most 65536 probes. */
typedef struct {
    double position[3];
    double direction[3]; /* static-observer direction at the end point */
    double drift;        /* max relative E / L drift over the path */
    uint64_t steps;
    uint32_t outcome;    /* GARG_RAY_CAPTURED / ESCAPED / EXHAUSTED */
    uint32_t reserved;
} garg_ray_probe_fp64_result;
int garg_probe_fp64(garg_ctx* ctx, garg_stream* stream, const garg_ray_probe* probes, uint32_t count,
                    const garg_geodesic_settings* settings, garg_ray_probe_fp64_result* results);

/* The reference Oracle's lensed render (Oracle/LensedRender.swift:46-91:
 * LensedRender.Camera, trace, image) on the GPU in double precision: every
 * pixel's camera ray integrated with the FP64 integrator of
 * garg_probe_fp64 and composited through the resolved volume, sampled as
 * VolumeField.sample does (VolumeField.swift:40-63) from the fp16 fields
 * widened to double. The reference's beaming golden compares the GPU march's
 * image with this render (Validate.swift:258-291); `gargantua validate`
 * judges it against this. `rgb` receives 3 doubles per pixel, row major,
 * linear and without background. Synchronous; at most 2^22 pixels. */
typedef struct {
    double position[3];
    double right[3];
    double up[3];
    double forward[3];
    double tan_half_fov[2]; /* tan(fov / 2) scaled by the aspect in x, and in y */
} garg_render_camera_fp64;
int garg_render_fp64(garg_ctx* ctx, garg_stream* stream, const garg_volume* volume,
                     const garg_render_camera_fp64* camera, uint32_t width, uint32_t height,
                     const garg_geodesic_settings* settings, double opacity, uint32_t redshift, double* rgb);

/* ---- measurement ------------------------------------------------------------ */

/* No counterpart in the reference: the ceiling the splat bin is measured
 * against. Times `threads` x `adds_per_thread` no-return 32-bit atomic adds
 * to pcg4d-chosen words of a `words`-word buffer (allocated and freed here)
 * and returns the rate. Synchronous. */
int garg_atomic_rate(garg_ctx* ctx, garg_stream* stream, uint32_t words, uint32_t adds_per_thread, uint32_t threads,
                     double* adds_per_second);

/* No counterpart in the reference: the FP32 ceilings the march is measured
 * against. Times `threads` x `iterations` x 32 independent non-packed FP32
 * fused multiply-adds (2 flops each) in one operand form and returns flops
 * per second. The form sets the issue rate: two VGPRs and a literal
 * (v_fmamk_f32), three VGPRs (v_fma_f32), or SGPR operands (wave-uniform
 * values). Synchronous. */
enum { GARG_VALU_FMA_LITERAL = 0, GARG_VALU_FMA_VGPR = 1, GARG_VALU_FMA_SGPR = 2 };
int garg_valu_rate(garg_ctx* ctx, garg_stream* stream, uint32_t form, uint32_t iterations, uint32_t threads,
                   double* flops_per_second);

/* ---- graphs --------------------------------------------------------------- */

/* HIP graph capture of a launch sequence, the replacement for re-encoding
 * the same command buffer every frame (Renderer.swift:119-215 encodes sim,
 * splat and resolve per frame). Between begin and end every launch entry
 * point enqueued on `stream` is recorded instead of run; entry points that
 * copy to or from the host or wait for the device (reads, checksums, table
 * updates, probes, the batched march, tonemap readback, syncs, destroys)
 * refuse with GARG_E_STATE while the stream or the context's stream
 * captures. No entry point records a memset node: clears are kernels.
 * A replay runs the recorded launches with the arguments they had at
 * capture (use GARG_FRAME_AUTO for a splat frame that advances). */
int garg_graph_begin(garg_ctx* ctx, garg_stream* stream);
int garg_graph_end(garg_ctx* ctx, garg_stream* stream, garg_graph** out);
int garg_graph_launch(garg_graph* graph, garg_stream* stream);
int garg_graph_destroy(garg_graph* graph);

#ifdef __cplusplus
}
#endif

#endif
