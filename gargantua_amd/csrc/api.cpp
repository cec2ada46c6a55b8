// The C ABI of libgargantua_hip.so (include/gargantua.h): objects, memory
// budget, argument checks and kernel launches. Every entry point returns a
// status and leaves one sentence in garg_last_error() on failure.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gargantua.h"
#include "../../include/gargantua_constants.h"
#include "host_tables.hpp"
#include "launch.hpp"

using namespace garg;

namespace {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define HIP_TRY(expr, what)                                                                 \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail(GARG_E_DEVICE, "%s: %s", what, hipGetErrorString(e_)); \
    } while (0)

#define REQUIRE(cond, ...)                                             \
    do {                                                               \
        if (!(cond)) return fail(GARG_E_ARGUMENT, __VA_ARGS__);        \
    } while (0)

double max_impact_parameter(double r) { return r / std::sqrt(1.0 - R_SCHWARZSCHILD / r); }

// Staging of garg_probe_noalloc: the uploaded probes, then the results, for at
// most GARG_PROBE_NOALLOC_MAX rays. Every device holds its own copy of this
// array until the process ends. In host code the name `scratch` is only a
// host-side stand-in: the device address of each device's copy is looked up
// once, with that device current, and kept in g_scratch_ptr[device]; host code
// passes that pointer, never the array's name, to kernels and copies.
constexpr uint64_t NOALLOC_PROBES_BYTES = ((uint64_t)GARG_PROBE_NOALLOC_MAX * sizeof(garg_ray_probe) + 255u) & ~255ull;
constexpr uint64_t NOALLOC_SCRATCH_BYTES =
    NOALLOC_PROBES_BYTES + (uint64_t)GARG_PROBE_NOALLOC_MAX * sizeof(garg_ray_probe_result);
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Looks up this device's copy of `scratch` on the first context created on
// it; `device` must be current.
int bind_noalloc_scratch(int device) {
    if (device >= MAX_DEVICES) return GARG_OK;  // garg_probe_noalloc refuses on such a device
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return GARG_OK;
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "locating the probe staging array");
    g_scratch_ptr[device] = p;
    return GARG_OK;
}

}  // namespace

struct garg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t allocated = 0;
    garg_device_info info{};
    float4* blackbody = nullptr;
    float* inverse_cdf = nullptr;
    double temperature_scale = 0.0;
    // Scratch of the batched march and the probes.
    // Two halves each (spinning pass, Schwarzschild pass), so both passes of
    // one call enqueue without waiting for each other.
    BatchRequest* batch_requests = nullptr;
    BatchRequest* batch_staging = nullptr;  // pinned
    hipEvent_t batch_staging_free[2] = {nullptr, nullptr};
    hipEvent_t batch_done = nullptr;        // the previous call's launches have run
    unsigned long long* queue = nullptr;
    garg_march_counters* batch_totals = nullptr;
    // Workgroups per compute unit of the one-launch march, per integrator
    // (0 = not queried yet), for garg_march_batch's lone-request choice.
    uint32_t image_resident[4] = {0, 0, 0, 0};
    garg_march_counters* lone_counters = nullptr;  // a lone request's counts in garg_march_batch
    void* probe_scratch = nullptr;
    uint64_t probe_scratch_bytes = 0;
};

struct garg_stream {
    garg_ctx* ctx;
    hipStream_t s;
};

struct garg_event {
    garg_ctx* ctx;
    hipEvent_t e;
};

struct garg_particles {
    garg_ctx* ctx;
    uint32_t count;
    uint64_t seed;
    uint32_t potential;
    void* base;
    uint64_t bytes;
    ParticleView view;
    void* sort_base;  // the exact deposit's keys, order, staged particles and sort storage
    uint64_t sort_bytes;
    SplatSortBuffers sort;
};

struct garg_volume {
    garg_ctx* ctx;
    uint32_t size;
    float half_xy, half_z;
    void* base;
    uint64_t bytes;
    uint32_t* emission;
    int32_t* momentum;
    uint32_t* mass;
    Voxel* voxels;
    void* gather_base;  // the exact deposit's cell starts and brick stamps
    uint64_t gather_bytes;
    SplatGather gather;
    void* bins_base;  // the one-corner deposit's brick tables (null past its largest side)
    uint64_t bins_bytes;
    SplatBins bins;
    uint32_t splat_method;  // GARG_SPLAT_SORTED or GARG_SPLAT_ATOMIC
};

struct garg_tables {
    garg_ctx* ctx;
    double sphere_radius, entry_radius, camera_radius;
    float* data;  // sphere, camera, sky
    float sphere_max_b, sphere_table_max_b, camera_max_b, sky_min_b, sky_max_b;
};

struct garg_targets {
    garg_ctx* ctx;
    uint32_t width, height;
    void* base;
    uint64_t bytes;
    h4* color;
    h4* debug;
    float* depth;
    garg_march_counters* counters;
    uint8_t* rgba8;
};

struct garg_bake {
    garg_ctx* ctx;
    uint32_t width, height, capacity;
    void* base;
    uint64_t bytes;
    h4* samples;
    uint4* headers;
    uint32_t* overflow;
};

namespace {

hipStream_t stream_of(garg_ctx* ctx, garg_stream* s) { return s ? s->s : ctx->stream; }

uint64_t align256(uint64_t b) { return (b + 255u) & ~(uint64_t)255u; }

int ctx_alloc(garg_ctx* ctx, uint64_t bytes, void** out, const char* what) {
    bytes = align256(std::max<uint64_t>(bytes, 256));
    if (ctx->allocated + bytes > ctx->budget)
        return fail(GARG_E_BUDGET, "%s needs %.1f MB but the context budget has %.1f MB of %.1f MB left.", what,
                    bytes / 1e6, (ctx->budget - ctx->allocated) / 1e6, ctx->budget / 1e6);
    HIP_TRY(hipMalloc(out, bytes), what);
    ctx->allocated += bytes;
    return GARG_OK;
}

void ctx_free(garg_ctx* ctx, void* p, uint64_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    ctx->allocated -= align256(std::max<uint64_t>(bytes, 256));
}

int sync_stream(hipStream_t s, const char* what) {
    HIP_TRY(hipStreamSynchronize(s), what);
    return GARG_OK;
}

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus status = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &status) == hipSuccess && status == hipStreamCaptureStatusActive;
}

// Entry points that copy to or from the host, or wait for the device,
// cannot be recorded into a graph: they refuse while either the stream they
// were given or the context's stream is capturing.
#define REFUSE_WHILE_CAPTURING(CTX_, STREAM_, what)                                                      \
    do {                                                                                                 \
        if (capturing((CTX_)->stream) || ((STREAM_) && capturing((STREAM_)->s)))                          \
            return fail(GARG_E_STATE, "%s copies to or from the host or waits for the device, so it cannot " \
                        "run while a graph is being captured.", what);                                    \
    } while (0)

Scene make_scene(garg_ctx* ctx, garg_volume* volume, garg_tables* tables) {
    Scene sc{};
    if (volume)
        sc.volume = VolumeView{volume->voxels, (int)volume->size, volume->half_xy, volume->half_z,
                               (float)volume->size / (2.0f * volume->half_xy), (float)volume->size / (2.0f * volume->half_z)};
    if (tables) {
        sc.sphere_table = tables->data;
        sc.camera_table = tables->data + SWEEP_TABLE_SIZE;
        sc.sky_table = tables->data + 2 * SWEEP_TABLE_SIZE;
    }
    sc.blackbody = ctx->blackbody;
    return sc;
}

int check_uniforms(const garg_march_uniforms* u) {
    REQUIRE(u, "uniforms are required.");
    REQUIRE(u->resolution[0] > 0 && u->resolution[1] > 0, "the resolution must be positive.");
    REQUIRE(u->spin >= 0.0f && u->spin < 1.0f, "spin must be in [0, 1), not %g.", (double)u->spin);
    REQUIRE(u->geodesic.step_cap <= (1u << 24), "the step cap %u is above 2^24.", u->geodesic.step_cap);
    REQUIRE(u->geodesic.escape_radius > u->geodesic.capture_radius, "the escape radius must exceed the capture radius.");
    float r = std::sqrt(u->camera_position[0] * u->camera_position[0] + u->camera_position[1] * u->camera_position[1] +
                        u->camera_position[2] * u->camera_position[2]);
    REQUIRE(std::isfinite(r) && r > 2.5f, "the camera must sit outside r = 2.5, not at r = %g.", (double)r);
    if (u->spin > 0.0f) {
        REQUIRE(u->geodesic.escape_radius <= (float)KERR_MAX_RADIUS && r <= (float)KERR_MAX_RADIUS,
                "spinning rays are marched within r = %g: escape radius %g, camera at r = %g.", KERR_MAX_RADIUS,
                (double)u->geodesic.escape_radius, (double)r);
    }
    return GARG_OK;
}

// Uniforms as the device reads them: a zero z extent means a cube.
garg_march_uniforms device_uniforms(const garg_march_uniforms* u, garg_volume* volume) {
    garg_march_uniforms d = *u;
    if (volume) {
        d.volume_half_extent = volume->half_xy;
        d.volume_half_extent_z = volume->half_z;
    } else if (d.volume_half_extent_z <= 0.0f) {
        d.volume_half_extent_z = d.volume_half_extent;
    }
    return d;
}

}  // namespace

extern "C" {

const char* garg_last_error(void) { return g_error.c_str(); }
const char* garg_version(void) { return "gargantua-mi355x 0.1.0 (gfx950)"; }

int garg_ctx_create(int device, uint64_t budget_bytes, garg_ctx** out) {
    REQUIRE(out, "out is required.");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count), "counting HIP devices");
    REQUIRE(count > 0, "no HIP device is visible.");
    REQUIRE(device >= 0 && device < count, "device %d does not exist; %d are visible.", device, count);
    HIP_TRY(hipSetDevice(device), "selecting the device");
    garg_ctx* ctx = new garg_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device), "reading device properties");
    snprintf(ctx->info.name, sizeof(ctx->info.name), "%s", prop.name);
    snprintf(ctx->info.arch, sizeof(ctx->info.arch), "%s", prop.gcnArchName);
    ctx->info.compute_units = (uint32_t)prop.multiProcessorCount;
    ctx->info.clock_khz = (uint32_t)prop.clockRate;
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b), "reading device memory");
    ctx->info.total_bytes = total_b;
    ctx->info.free_bytes = free_b;
    ctx->budget = budget_bytes ? budget_bytes : (uint64_t)(0.9 * (double)free_b);
    ctx->info.budget_bytes = ctx->budget;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "creating the default stream");
    int rc = bind_noalloc_scratch(device);
    if (rc) return rc;

    std::vector<float> bb, cdf;
    build_blackbody_table(bb);
    build_inverse_cdf(cdf);
    ctx->temperature_scale = temperature_scale();
    if ((rc = ctx_alloc(ctx, bb.size() * 4, (void**)&ctx->blackbody, "the blackbody table"))) return rc;
    if ((rc = ctx_alloc(ctx, cdf.size() * 4, (void**)&ctx->inverse_cdf, "the seeding table"))) return rc;
    HIP_TRY(hipMemcpy(ctx->blackbody, bb.data(), bb.size() * 4, hipMemcpyHostToDevice), "uploading the blackbody table");
    HIP_TRY(hipMemcpy(ctx->inverse_cdf, cdf.data(), cdf.size() * 4, hipMemcpyHostToDevice), "uploading the seeding table");
    if ((rc = ctx_alloc(ctx, 2 * sizeof(BatchRequest) * GARG_BATCH_MAX_REQUESTS, (void**)&ctx->batch_requests,
                        "the batch request table")))
        return rc;
    HIP_TRY(hipHostMalloc((void**)&ctx->batch_staging, 2 * sizeof(BatchRequest) * GARG_BATCH_MAX_REQUESTS, 0),
            "allocating pinned staging");
    for (hipEvent_t& e : ctx->batch_staging_free)
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "creating an event");
    HIP_TRY(hipEventCreateWithFlags(&ctx->batch_done, hipEventDisableTiming), "creating an event");
    if ((rc = ctx_alloc(ctx, 2 * 8 * 16 * sizeof(unsigned long long), (void**)&ctx->queue, "the ray queue"))) return rc;
    if ((rc = ctx_alloc(ctx, sizeof(garg_march_counters), (void**)&ctx->batch_totals, "batch counters"))) return rc;
    if ((rc = ctx_alloc(ctx, sizeof(garg_march_counters), (void**)&ctx->lone_counters, "counters"))) return rc;
    *out = ctx;
    return GARG_OK;
}

int garg_ctx_destroy(garg_ctx* ctx) {
    if (!ctx) return GARG_OK;
    REFUSE_WHILE_CAPTURING(ctx, (garg_stream*)nullptr, "garg_ctx_destroy");
    (void)hipStreamSynchronize(ctx->stream);
    ctx_free(ctx, ctx->blackbody, BLACKBODY_TABLE_SIZE * 16);
    ctx_free(ctx, ctx->inverse_cdf, INVERSE_CDF_COUNT * 4);
    ctx_free(ctx, ctx->batch_requests, 2 * sizeof(BatchRequest) * GARG_BATCH_MAX_REQUESTS);
    ctx_free(ctx, ctx->queue, 2 * 8 * 16 * sizeof(unsigned long long));
    ctx_free(ctx, ctx->batch_totals, sizeof(garg_march_counters));
    ctx_free(ctx, ctx->lone_counters, sizeof(garg_march_counters));
    ctx_free(ctx, ctx->probe_scratch, ctx->probe_scratch_bytes);
    if (ctx->batch_staging) (void)hipHostFree(ctx->batch_staging);
    for (hipEvent_t e : ctx->batch_staging_free)
        if (e) (void)hipEventDestroy(e);
    if (ctx->batch_done) (void)hipEventDestroy(ctx->batch_done);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GARG_OK;
}

int garg_device_info_get(garg_ctx* ctx, garg_device_info* out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b), "reading device memory");
    ctx->info.free_bytes = free_b;
    *out = ctx->info;
    return GARG_OK;
}

int garg_ctx_allocated(garg_ctx* ctx, uint64_t* bytes) {
    REQUIRE(ctx && bytes, "ctx and bytes are required.");
    *bytes = ctx->allocated;
    return GARG_OK;
}

int garg_stream_create(garg_ctx* ctx, garg_stream** out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    hipStream_t s;
    // Default priority: every stream of the library runs at one priority.
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "creating a stream");
    *out = new garg_stream{ctx, s};
    return GARG_OK;
}

int garg_stream_destroy(garg_stream* stream) {
    if (!stream) return GARG_OK;
    (void)hipStreamSynchronize(stream->s);
    (void)hipStreamDestroy(stream->s);
    delete stream;
    return GARG_OK;
}

int garg_stream_sync(garg_ctx* ctx, garg_stream* stream) {
    REQUIRE(ctx, "ctx is required.");
    REFUSE_WHILE_CAPTURING(ctx, stream, "garg_stream_sync");
    return sync_stream(stream_of(ctx, stream), "waiting for the stream");
}

int garg_event_create(garg_ctx* ctx, garg_event** out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e), "creating an event");
    *out = new garg_event{ctx, e};
    return GARG_OK;
}

int garg_event_destroy(garg_event* event) {
    if (!event) return GARG_OK;
    (void)hipEventDestroy(event->e);
    delete event;
    return GARG_OK;
}

int garg_event_record(garg_event* event, garg_stream* stream) {
    REQUIRE(event, "event is required.");
    HIP_TRY(hipEventRecord(event->e, stream_of(event->ctx, stream)), "recording an event");
    return GARG_OK;
}

int garg_event_sync(garg_event* event) {
    REQUIRE(event, "event is required.");
    HIP_TRY(hipEventSynchronize(event->e), "waiting for an event");
    return GARG_OK;
}

int garg_event_elapsed_ms(garg_event* start, garg_event* end, float* ms) {
    REQUIRE(start && end && ms, "start, end and ms are required.");
    HIP_TRY(hipEventElapsedTime(ms, start->e, end->e), "reading elapsed time");
    return GARG_OK;
}

// ---- particles -------------------------------------------------------------

int garg_particles_create(garg_ctx* ctx, uint32_t count, uint64_t seed, uint32_t potential, garg_particles** out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    REQUIRE(count > 0, "the particle count must be positive.");
    REQUIRE(potential <= 1u, "unknown potential %u.", potential);
    garg_particles* p = new garg_particles();
    p->ctx = ctx;
    p->count = count;
    p->seed = seed;
    p->potential = potential;
    uint64_t field = align256((uint64_t)count * 4u);
    p->bytes = field * 8u + 256u;
    int rc = ctx_alloc(ctx, p->bytes, &p->base, "the particle store");
    if (rc) {
        delete p;
        return rc;
    }
    char* b = static_cast<char*>(p->base);
    p->view = ParticleView{(float*)(b), (float*)(b + field), (float*)(b + 2 * field), (float*)(b + 3 * field),
                           (float*)(b + 4 * field), (float*)(b + 5 * field), (float*)(b + 6 * field),
                           (uint32_t*)(b + 7 * field), (uint32_t*)(b + 8 * field), count};
    // The dispatch count starts at zero even before the first seed.
    rc = launch_fill_zero(p->view.dispatches, 256, ctx->stream) == hipSuccess ? GARG_OK
                                                                               : fail(GARG_E_DEVICE, "clearing the dispatch count");
    if (!rc) {
        uint64_t words = align256((uint64_t)count * 4u);
        p->sort.temp_bytes = splat_sort_temp_bytes(count);
        p->sort_bytes = 20u * words + align256(p->sort.temp_bytes);
        rc = ctx_alloc(ctx, p->sort_bytes, &p->sort_base, "the deposit's sort buffers");
        if (!rc) {
            char* s = static_cast<char*>(p->sort_base);
            p->sort.keys_in = (uint32_t*)s;
            p->sort.keys_out = (uint32_t*)(s + words);
            p->sort.order_in = (uint32_t*)(s + 2 * words);
            p->sort.order_out = (uint32_t*)(s + 3 * words);
            p->sort.records = (float4*)(s + 4 * words);
            p->sort.sorted = (float4*)(s + 12 * words);
            p->sort.temp = s + 20 * words;
        }
    }
    if (rc) {
        ctx_free(ctx, p->base, p->bytes);
        delete p;
        return rc;
    }
    *out = p;
    return GARG_OK;
}

int garg_particles_destroy(garg_particles* p) {
    if (!p) return GARG_OK;
    REFUSE_WHILE_CAPTURING(p->ctx, (garg_stream*)nullptr, "garg_particles_destroy");
    (void)hipStreamSynchronize(p->ctx->stream);
    ctx_free(p->ctx, p->base, p->bytes);
    ctx_free(p->ctx, p->sort_base, p->sort_bytes);
    delete p;
    return GARG_OK;
}

static SimArgs sim_args(garg_particles* p, uint32_t generator, float sigma_z) {
    SimArgs a;
    a.p = p->view;
    a.seed_lo = (uint32_t)p->seed;
    a.seed_hi = (uint32_t)(p->seed >> 32);
    a.potential = p->potential;
    a.generator = generator;
    a.sigma_z = sigma_z;
    a.dt = GARG_F(SIM_DT);
    a.alpha = GARG_F(DRAG_ALPHA);
    a.temperature_scale = (float)p->ctx->temperature_scale;
    a.inverse_cdf = p->ctx->inverse_cdf;
    a.inverse_cdf_count = INVERSE_CDF_COUNT;
    return a;
}

int garg_particles_seed(garg_particles* p, garg_stream* stream, uint32_t generator, float sigma_z) {
    REQUIRE(p, "particles are required.");
    REQUIRE(generator <= 1u, "unknown generator %u.", generator);
    REQUIRE(generator == GARG_GENERATOR_DISK || (sigma_z >= 0.0f && std::isfinite(sigma_z)),
            "sigma_z must be a finite non-negative height.");
    HIP_TRY(launch_seed(sim_args(p, generator, sigma_z), stream_of(p->ctx, stream)), "launching seedParticles");
    return GARG_OK;
}

int garg_particles_step(garg_particles* p, garg_stream* stream, uint32_t dispatches) {
    REQUIRE(p, "particles are required.");
    SimArgs a = sim_args(p, GARG_GENERATOR_DISK, 0.0f);
    hipStream_t s = stream_of(p->ctx, stream);
    // Up to 256 dispatches per launch (each launch stays a few milliseconds
    // at the largest stores).
    for (uint32_t done = 0; done < dispatches;) {
        const uint32_t n = std::min(dispatches - done, 256u);
        HIP_TRY(launch_step(a, n, s), "launching stepParticles");
        done += n;
    }
    return GARG_OK;
}

int garg_particles_reseed(garg_particles* p, uint64_t seed) {
    REQUIRE(p, "particles are required.");
    p->seed = seed;
    return GARG_OK;
}

int garg_particles_read(garg_particles* p, garg_particle* host, uint32_t count) {
    REQUIRE(p && host, "particles and host are required.");
    REFUSE_WHILE_CAPTURING(p->ctx, (garg_stream*)nullptr, "garg_particles_read");
    REQUIRE(count == p->count, "the store holds %u particles, not %u.", p->count, count);
    garg_particle* tmp = nullptr;
    int rc = ctx_alloc(p->ctx, (uint64_t)count * sizeof(garg_particle), (void**)&tmp, "the particle readback");
    if (rc) return rc;
    hipStream_t s = p->ctx->stream;
    hipError_t e = launch_pack_aos(p->view, tmp, s);
    if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, (size_t)count * sizeof(garg_particle), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ctx_free(p->ctx, tmp, (uint64_t)count * sizeof(garg_particle));
    HIP_TRY(e, "reading the particles back");
    return GARG_OK;
}

int garg_particles_write(garg_particles* p, const garg_particle* host, uint32_t count) {
    REQUIRE(p && host, "particles and host are required.");
    REFUSE_WHILE_CAPTURING(p->ctx, (garg_stream*)nullptr, "garg_particles_write");
    REQUIRE(count == p->count, "the store holds %u particles, not %u.", p->count, count);
    garg_particle* tmp = nullptr;
    int rc = ctx_alloc(p->ctx, (uint64_t)count * sizeof(garg_particle), (void**)&tmp, "the particle upload");
    if (rc) return rc;
    hipStream_t s = p->ctx->stream;
    hipError_t e = hipMemcpyAsync(tmp, host, (size_t)count * sizeof(garg_particle), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_unpack_aos(p->view, tmp, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ctx_free(p->ctx, tmp, (uint64_t)count * sizeof(garg_particle));
    HIP_TRY(e, "writing the particles");
    return GARG_OK;
}

int garg_particles_checksum(garg_particles* p, uint64_t* fnv1a) {
    REQUIRE(p && fnv1a, "particles and fnv1a are required.");
    REFUSE_WHILE_CAPTURING(p->ctx, (garg_stream*)nullptr, "garg_particles_checksum");
    std::vector<garg_particle> host(p->count);
    int rc = garg_particles_read(p, host.data(), p->count);
    if (rc) return rc;
    const uint8_t* bytes = reinterpret_cast<const uint8_t*>(host.data());
    uint64_t hash = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < host.size() * sizeof(garg_particle); ++i) {
        hash ^= bytes[i];
        hash *= 0x100000001b3ull;
    }
    *fnv1a = hash;
    return GARG_OK;
}

// ---- volume ----------------------------------------------------------------

int garg_volume_create(garg_ctx* ctx, uint32_t size, float half_extent, float half_extent_z, garg_volume** out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    REQUIRE(size >= 2 && size <= 1024, "the volume size must be in [2, 1024], not %u.", size);
    REQUIRE(half_extent > 0.0f && half_extent_z >= 0.0f, "half extents must be positive.");
    garg_volume* v = new garg_volume();
    v->ctx = ctx;
    v->size = size;
    v->half_xy = half_extent;
    v->half_z = half_extent_z > 0.0f ? half_extent_z : half_extent;
    uint64_t voxels = (uint64_t)size * size * size;
    uint64_t sums = align256(voxels * 24u), fields = align256(voxels * 16u);
    v->bytes = sums + fields;
    int rc = ctx_alloc(ctx, v->bytes, &v->base, "the volume");
    if (rc) {
        delete v;
        return rc;
    }
    char* b = static_cast<char*>(v->base);
    v->emission = (uint32_t*)b;
    v->momentum = (int32_t*)(b + voxels * 8u);
    v->mass = (uint32_t*)(b + voxels * 20u);
    v->voxels = (Voxel*)(b + sums);
    const uint64_t m = (uint64_t)size + 1u;
    const uint64_t cells = align256(m * m * m * 4u);
    v->gather_bytes = 2u * cells + align256(splat_gather_bricks_max(size) * 4u);
    rc = ctx_alloc(ctx, v->gather_bytes, &v->gather_base, "the deposit's cell tables");
    if (rc) {
        ctx_free(ctx, v->base, v->bytes);
        delete v;
        return rc;
    }
    const uint64_t bin_words = splat_bin_table_words(size);
    v->bins = splat_bin_layout(size);
    v->bins_base = nullptr;
    v->bins_bytes = align256(bin_words * 4u);
    if (bin_words) {
        rc = ctx_alloc(ctx, v->bins_bytes, &v->bins_base, "the deposit's brick tables");
        if (rc) {
            ctx_free(ctx, v->base, v->bytes);
            ctx_free(ctx, v->gather_base, v->gather_bytes);
            delete v;
            return rc;
        }
        v->bins.counts = (uint32_t*)v->bins_base;
        v->bins.totals = v->bins.counts + (uint64_t)kSplatBinGroups * v->bins.bricks;
        v->bins.starts = v->bins.totals + v->bins.bricks;
    }
    v->splat_method = GARG_SPLAT_SORTED;
    v->gather.cells = (uint32_t*)v->gather_base;
    v->gather.ends = (uint32_t*)((char*)v->gather_base + cells);
    v->gather.bricks = (uint32_t*)((char*)v->gather_base + 2u * cells);
    v->gather.stamp = 0;
    for (int k = 0; k < 3; ++k) v->gather.brick[k] = kSplatBricks[0][k];
    hipError_t e = launch_fill_zero(v->base, v->bytes, ctx->stream);
    if (e == hipSuccess) e = launch_fill_zero(v->gather.bricks, v->gather_bytes - 2u * cells, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx_free(ctx, v->base, v->bytes);
        ctx_free(ctx, v->gather_base, v->gather_bytes);
        if (v->bins_base) ctx_free(ctx, v->bins_base, v->bins_bytes);
        delete v;
        return fail(GARG_E_DEVICE, "clearing the volume: %s", hipGetErrorString(e));
    }
    *out = v;
    return GARG_OK;
}

int garg_volume_destroy(garg_volume* v) {
    if (!v) return GARG_OK;
    REFUSE_WHILE_CAPTURING(v->ctx, (garg_stream*)nullptr, "garg_volume_destroy");
    (void)hipStreamSynchronize(v->ctx->stream);
    ctx_free(v->ctx, v->base, v->bytes);
    ctx_free(v->ctx, v->gather_base, v->gather_bytes);
    if (v->bins_base) ctx_free(v->ctx, v->bins_base, v->bins_bytes);
    delete v;
    return GARG_OK;
}

int garg_volume_clear(garg_volume* v, garg_stream* stream) {
    REQUIRE(v, "volume is required.");
    uint64_t voxels = (uint64_t)v->size * v->size * v->size;
    HIP_TRY(launch_fill_zero(v->base, voxels * 24u, stream_of(v->ctx, stream)), "clearing the sums");
    return GARG_OK;
}

int garg_splat_bin(garg_volume* v, garg_particles* p, garg_stream* stream, uint32_t exact, uint32_t frame) {
    REQUIRE(v && p, "volume and particles are required.");
    REQUIRE(v->ctx == p->ctx, "volume and particles belong to different contexts.");
    SplatArgs a;
    a.p = p->view;
    a.emission = v->emission;
    a.momentum = v->momentum;
    a.mass = v->mass;
    a.size = v->size;
    a.half_xy = v->half_xy;
    a.half_z = v->half_z;
    a.peak_temperature = GARG_F(T_PEAK_DEFAULT);
    a.emission_scale = GARG_F(EMISSION_SCALE);
    a.exact = exact ? 1u : 0u;
    a.frame = frame;
    a.seed_lo = (uint32_t)p->seed;
    a.seed_hi = (uint32_t)(p->seed >> 32);
    a.frame_counter = frame == GARG_FRAME_AUTO ? p->view.dispatches : nullptr;
    hipStream_t s = stream_of(v->ctx, stream);
    // The sorted deposits give the atomic deposit's integer sums without its
    // device-scope atomics. The exact one sorts with the library's radix
    // sort, which clears its storage with memset nodes that graphs must not
    // hold, so inside a capture it deposits directly; the one-corner deposit
    // sorts by brick with kernels only, captured or not, on volumes whose
    // bricks its tables hold.
    const bool sorted = v->splat_method == GARG_SPLAT_SORTED;
    if (sorted && a.exact && !capturing(s)) {
        // A new stamp per deposit; zero is the cleared bricks' value.
        if (++v->gather.stamp == 0u) v->gather.stamp = 1u;
        HIP_TRY(launch_splat_sorted(a, p->sort, v->gather, s), "launching the sorted splat");
    } else if (sorted && !a.exact && v->bins.bricks != 0u) {
        SplatBins bins = v->bins;
        bins.keys = p->sort.keys_in;
        bins.records = reinterpret_cast<uint4*>(p->sort.records);
        HIP_TRY(launch_splat_binned(a, bins, s), "launching the binned splat");
    } else {
        HIP_TRY(launch_splat_bin(a, s), "launching splatParticles");
    }
    return GARG_OK;
}

int garg_volume_set_splat_method(garg_volume* v, uint32_t method) {
    REQUIRE(v, "volume is required.");
    REQUIRE(method <= GARG_SPLAT_ATOMIC, "unknown splat method %u (0 sorted, 1 atomic).", method);
    v->splat_method = method;
    return GARG_OK;
}

int garg_volume_splat_method(garg_volume* v, uint32_t* method) {
    REQUIRE(v && method, "volume and method are required.");
    *method = v->splat_method;
    return GARG_OK;
}

int garg_volume_set_splat_brick(garg_volume* v, uint32_t x, uint32_t y, uint32_t z) {
    REQUIRE(v, "volume is required.");
    const uint32_t shape[3] = {x, y, z};
    REQUIRE(splat_brick_supported(shape),
            "brick %ux%ux%u is not one of the exact deposit's shapes (16x8x4, 8x8x4, 8x8x8, 16x16x4, 32x8x4, "
            "16x8x8, 32x16x4, 16x16x8).",
            x, y, z);
    for (int k = 0; k < 3; ++k) v->gather.brick[k] = shape[k];
    return GARG_OK;
}

int garg_volume_splat_brick(garg_volume* v, uint32_t shape[3]) {
    REQUIRE(v && shape, "volume and shape are required.");
    for (int k = 0; k < 3; ++k) shape[k] = v->gather.brick[k];
    return GARG_OK;
}

int garg_splat_resolve(garg_volume* v, garg_stream* stream, uint32_t clear_sums) {
    REQUIRE(v, "volume is required.");
    ResolveArgs a;
    a.emission = v->emission;
    a.momentum = v->momentum;
    a.mass = v->mass;
    a.voxels = v->voxels;
    a.blackbody = v->ctx->blackbody;
    a.size = v->size;
    a.half_xy = v->half_xy;
    a.half_z = v->half_z;
    a.peak_temperature = GARG_F(T_PEAK_DEFAULT);
    a.clear = clear_sums ? 1u : 0u;
    HIP_TRY(launch_resolve(a, stream_of(v->ctx, stream)), "launching resolveVolume");
    return GARG_OK;
}

int garg_volume_read_sums(garg_volume* v, uint32_t* emission, int32_t* momentum, uint32_t* mass) {
    REQUIRE(v, "volume is required.");
    REFUSE_WHILE_CAPTURING(v->ctx, (garg_stream*)nullptr, "garg_volume_read_sums");
    uint64_t voxels = (uint64_t)v->size * v->size * v->size;
    hipStream_t s = v->ctx->stream;
    if (emission) HIP_TRY(hipMemcpyAsync(emission, v->emission, voxels * 8u, hipMemcpyDeviceToHost, s), "reading sums");
    if (momentum) HIP_TRY(hipMemcpyAsync(momentum, v->momentum, voxels * 12u, hipMemcpyDeviceToHost, s), "reading sums");
    if (mass) HIP_TRY(hipMemcpyAsync(mass, v->mass, voxels * 4u, hipMemcpyDeviceToHost, s), "reading sums");
    return sync_stream(s, "reading sums");
}

int garg_volume_read_fields(garg_volume* v, uint16_t* emission, uint16_t* velocity) {
    REQUIRE(v, "volume is required.");
    REFUSE_WHILE_CAPTURING(v->ctx, (garg_stream*)nullptr, "garg_volume_read_fields");
    uint64_t voxels = (uint64_t)v->size * v->size * v->size;
    hipStream_t s = v->ctx->stream;
    if (emission) HIP_TRY(hipMemcpy2DAsync(emission, 8, v->voxels, 16, 8, voxels, hipMemcpyDeviceToHost, s), "reading fields");
    if (velocity)
        HIP_TRY(hipMemcpy2DAsync(velocity, 8, (char*)v->voxels + 8, 16, 8, voxels, hipMemcpyDeviceToHost, s),
                "reading fields");
    return sync_stream(s, "reading fields");
}

int garg_volume_write_fields(garg_volume* v, const uint16_t* emission, const uint16_t* velocity) {
    REQUIRE(v && emission && velocity, "volume, emission and velocity are required.");
    REFUSE_WHILE_CAPTURING(v->ctx, (garg_stream*)nullptr, "garg_volume_write_fields");
    uint64_t voxels = (uint64_t)v->size * v->size * v->size;
    hipStream_t s = v->ctx->stream;
    HIP_TRY(hipMemcpy2DAsync(v->voxels, 16, emission, 8, 8, voxels, hipMemcpyHostToDevice, s), "writing fields");
    HIP_TRY(hipMemcpy2DAsync((char*)v->voxels + 8, 16, velocity, 8, 8, voxels, hipMemcpyHostToDevice, s),
            "writing fields");
    return sync_stream(s, "writing fields");
}

// ---- sweep tables ------------------------------------------------------------

int garg_tables_create(garg_ctx* ctx, garg_stream* stream, float sphere_radius, garg_tables** out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    REFUSE_WHILE_CAPTURING(ctx, stream, "garg_tables_create");
    REQUIRE(sphere_radius > 4.0f, "the integration sphere must lie outside r = 4.");
    garg_tables* t = new garg_tables();
    t->ctx = ctx;
    t->sphere_radius = sphere_radius;
    // Rays whose periapsis lies outside r = R - 1 never enter, so no ray is
    // placed on the sphere grazing it (DeflectionTables.swift:24-31).
    t->entry_radius = (double)sphere_radius - 1.0;
    t->camera_radius = 0.0;
    int rc = ctx_alloc(ctx, 3u * SWEEP_TABLE_SIZE * 4u, (void**)&t->data, "the sweep tables");
    if (rc) {
        delete t;
        return rc;
    }
    t->sphere_max_b = (float)max_impact_parameter(t->entry_radius);
    t->sky_min_b = t->sphere_max_b;
    t->sphere_table_max_b = (float)max_impact_parameter(t->sphere_radius);
    HIP_TRY(launch_sweep_table(t->data, SWEEP_TABLE_SIZE, t->sphere_radius, 0.0, max_impact_parameter(t->sphere_radius),
                               SWEEP_INTERVALS_SPHERE, 0, stream_of(ctx, stream)),
            "building the sphere sweep table");
    *out = t;
    return GARG_OK;
}

int garg_tables_destroy(garg_tables* t) {
    if (!t) return GARG_OK;
    REFUSE_WHILE_CAPTURING(t->ctx, (garg_stream*)nullptr, "garg_tables_destroy");
    (void)hipStreamSynchronize(t->ctx->stream);
    ctx_free(t->ctx, t->data, 3u * SWEEP_TABLE_SIZE * 4u);
    delete t;
    return GARG_OK;
}

int garg_tables_update(garg_tables* t, garg_stream* stream, float camera_radius) {
    REQUIRE(t, "tables are required.");
    REFUSE_WHILE_CAPTURING(t->ctx, stream, "garg_tables_update");
    REQUIRE(camera_radius > 3.5f, "the camera radius must exceed 3.5.");
    double r = (double)camera_radius;
    if (std::fabs(r - t->camera_radius) <= 1e-6) return GARG_OK;
    t->camera_radius = r;
    double bmax = max_impact_parameter(r);
    t->camera_max_b = (float)bmax;
    t->sky_max_b = t->camera_max_b;
    hipStream_t s = stream_of(t->ctx, stream);
    HIP_TRY(launch_sweep_table(t->data + SWEEP_TABLE_SIZE, SWEEP_TABLE_SIZE, r, 0.0, bmax, SWEEP_INTERVALS_CAMERA, 0, s),
            "building the camera sweep table");
    HIP_TRY(launch_sweep_table(t->data + 2 * SWEEP_TABLE_SIZE, SWEEP_TABLE_SIZE, r,
                               max_impact_parameter(t->entry_radius), bmax, SWEEP_INTERVALS_CAMERA, 1, s),
            "building the sky sweep table");
    return GARG_OK;
}

int garg_tables_apply(garg_tables* t, garg_march_uniforms* u) {
    REQUIRE(t && u, "tables and uniforms are required.");
    REQUIRE(t->camera_radius > 0.0, "the tables have no camera radius yet; call garg_tables_update first.");
    u->sphere_radius = (float)t->sphere_radius;
    u->camera_radius = (float)t->camera_radius;
    u->sphere_max_b = t->sphere_max_b;
    u->sphere_table_max_b = t->sphere_table_max_b;
    u->camera_max_b = t->camera_max_b;
    u->sky_min_b = t->sky_min_b;
    u->sky_max_b = t->sky_max_b;
    return GARG_OK;
}

int garg_tables_read(garg_tables* t, float* sphere, float* camera, float* sky) {
    REQUIRE(t, "tables are required.");
    REFUSE_WHILE_CAPTURING(t->ctx, (garg_stream*)nullptr, "garg_tables_read");
    hipStream_t s = t->ctx->stream;
    HIP_TRY(hipDeviceSynchronize(), "waiting for the device");
    if (sphere) HIP_TRY(hipMemcpyAsync(sphere, t->data, SWEEP_TABLE_SIZE * 4u, hipMemcpyDeviceToHost, s), "reading tables");
    if (camera)
        HIP_TRY(hipMemcpyAsync(camera, t->data + SWEEP_TABLE_SIZE, SWEEP_TABLE_SIZE * 4u, hipMemcpyDeviceToHost, s),
                "reading tables");
    if (sky)
        HIP_TRY(hipMemcpyAsync(sky, t->data + 2 * SWEEP_TABLE_SIZE, SWEEP_TABLE_SIZE * 4u, hipMemcpyDeviceToHost, s),
                "reading tables");
    return sync_stream(s, "reading tables");
}

// ---- targets ---------------------------------------------------------------

int garg_targets_create(garg_ctx* ctx, uint32_t width, uint32_t height, garg_targets** out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    REQUIRE(width > 0 && height > 0 && (uint64_t)width * height <= (1ull << 28), "bad target size %ux%u.", width, height);
    garg_targets* t = new garg_targets();
    t->ctx = ctx;
    t->width = width;
    t->height = height;
    uint64_t px = (uint64_t)width * height;
    uint64_t c = align256(px * 8u), dbg = align256(px * 8u), dep = align256(px * 4u), cnt = 256, rgba = align256(px * 4u);
    t->bytes = c + dbg + dep + cnt + rgba;
    int rc = ctx_alloc(ctx, t->bytes, &t->base, "the march targets");
    if (rc) {
        delete t;
        return rc;
    }
    char* b = static_cast<char*>(t->base);
    t->color = (h4*)b;
    t->debug = (h4*)(b + c);
    t->depth = (float*)(b + c + dbg);
    t->counters = (garg_march_counters*)(b + c + dbg + dep);
    t->rgba8 = (uint8_t*)(b + c + dbg + dep + cnt);
    hipError_t e = launch_fill_zero(t->base, t->bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    HIP_TRY(e, "clearing the targets");
    *out = t;
    return GARG_OK;
}

int garg_targets_destroy(garg_targets* t) {
    if (!t) return GARG_OK;
    REFUSE_WHILE_CAPTURING(t->ctx, (garg_stream*)nullptr, "garg_targets_destroy");
    (void)hipStreamSynchronize(t->ctx->stream);
    ctx_free(t->ctx, t->base, t->bytes);
    delete t;
    return GARG_OK;
}

int garg_targets_read(garg_targets* t, uint16_t* color, uint16_t* debug, float* depth) {
    REQUIRE(t, "targets are required.");
    REFUSE_WHILE_CAPTURING(t->ctx, (garg_stream*)nullptr, "garg_targets_read");
    HIP_TRY(hipDeviceSynchronize(), "waiting for the device");
    uint64_t px = (uint64_t)t->width * t->height;
    if (color) HIP_TRY(hipMemcpy(color, t->color, px * 8u, hipMemcpyDeviceToHost), "reading the colour target");
    if (debug) HIP_TRY(hipMemcpy(debug, t->debug, px * 8u, hipMemcpyDeviceToHost), "reading the debug target");
    if (depth) HIP_TRY(hipMemcpy(depth, t->depth, px * 4u, hipMemcpyDeviceToHost), "reading the depth target");
    return GARG_OK;
}

int garg_targets_read_counters(garg_targets* t, garg_march_counters* out) {
    REQUIRE(t && out, "targets and out are required.");
    REFUSE_WHILE_CAPTURING(t->ctx, (garg_stream*)nullptr, "garg_targets_read_counters");
    HIP_TRY(hipDeviceSynchronize(), "waiting for the device");
    HIP_TRY(hipMemcpy(out, t->counters, sizeof(garg_march_counters), hipMemcpyDeviceToHost), "reading counters");
    return GARG_OK;
}

int garg_targets_reset_counters(garg_targets* t, garg_stream* stream) {
    REQUIRE(t, "targets are required.");
    HIP_TRY(launch_fill_zero(t->counters, 256, stream_of(t->ctx, stream)), "clearing counters");
    return GARG_OK;
}

// ---- march -------------------------------------------------------------------

int garg_march_image(garg_ctx* ctx, garg_stream* stream, const garg_march_uniforms* uniforms, garg_volume* volume,
                     garg_tables* tables, garg_targets* targets, uint32_t extent_w, uint32_t extent_h,
                     uint32_t integrator) {
    REQUIRE(ctx && targets, "ctx and targets are required.");
    int rc = check_uniforms(uniforms);
    if (rc) return rc;
    REQUIRE(integrator <= 3u, "unknown integrator %u.", integrator);
    REQUIRE(targets->width == uniforms->resolution[0] && targets->height == uniforms->resolution[1],
            "the targets are %ux%u but the uniforms ask for %ux%u.", targets->width, targets->height,
            uniforms->resolution[0], uniforms->resolution[1]);
    REQUIRE(uniforms->tile_origin[0] < uniforms->resolution[0] && uniforms->tile_origin[1] < uniforms->resolution[1],
            "the tile origin lies outside the image.");
    REQUIRE(uniforms->spin > 0.0f || uniforms->sphere_radius <= 0.0f || tables,
            "the uniforms enable the sphere jump but no tables were given.");
    MarchImageArgs a;
    a.uniforms = device_uniforms(uniforms, volume);
    a.scene = make_scene(ctx, volume, tables);
    a.color = targets->color;
    a.debug = targets->debug;
    a.depth = targets->depth;
    a.counters = targets->counters;
    a.extent_w = extent_w ? std::min(extent_w, uniforms->resolution[0] - uniforms->tile_origin[0])
                          : uniforms->resolution[0] - uniforms->tile_origin[0];
    a.extent_h = extent_h ? std::min(extent_h, uniforms->resolution[1] - uniforms->tile_origin[1])
                          : uniforms->resolution[1] - uniforms->tile_origin[1];
    HIP_TRY(launch_march_image(a, integrator, stream_of(ctx, stream)), "launching marchImage");
    return GARG_OK;
}

namespace {

int check_probes(garg_ctx* ctx, garg_stream* stream, const garg_ray_probe* probes, uint32_t count,
                 const garg_march_uniforms* uniforms, garg_tables* tables, uint32_t integrator,
                 garg_ray_probe_result* results, const char* what) {
    REQUIRE(ctx && probes && results, "ctx, probes and results are required.");
    REFUSE_WHILE_CAPTURING(ctx, stream, what);
    REQUIRE(count > 0, "no probes.");
    REQUIRE(integrator <= 3u, "unknown integrator %u.", integrator);
    REQUIRE(uniforms && uniforms->spin >= 0.0f && uniforms->spin < 1.0f, "spin must be in [0, 1).");
    REQUIRE(uniforms->geodesic.step_cap <= (1u << 24), "the step cap is above 2^24.");
    REQUIRE(uniforms->spin > 0.0f || uniforms->sphere_radius <= 0.0f || tables,
            "the uniforms enable the sphere jump but no tables were given.");
    if (uniforms->spin > 0.0f) {
        REQUIRE(uniforms->geodesic.escape_radius <= (float)KERR_MAX_RADIUS,
                "spinning rays are marched within r = %g, not to an escape radius of %g.", KERR_MAX_RADIUS,
                (double)uniforms->geodesic.escape_radius);
        for (uint32_t i = 0; i < count; ++i) {
            const float* o = probes[i].origin;
            float r = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
            REQUIRE(std::isfinite(r) && r <= (float)KERR_MAX_RADIUS,
                    "spinning rays are marched within r = %g; probe %u starts at r = %g.", KERR_MAX_RADIUS, i, (double)r);
        }
    }
    return GARG_OK;
}

// Uploads the probes to `staging`, marches them and reads the results back
// from `staging` + results_offset; waits for the stream.
int run_probes(garg_ctx* ctx, garg_stream* stream, const garg_ray_probe* probes, uint32_t count,
               const garg_march_uniforms* uniforms, garg_tables* tables, uint32_t integrator,
               garg_ray_probe_result* results, void* staging, uint64_t results_offset) {
    hipStream_t s = stream_of(ctx, stream);
    ProbeArgs a;
    a.uniforms = device_uniforms(uniforms, nullptr);
    a.scene = make_scene(ctx, nullptr, tables);
    a.probes = (const garg_ray_probe*)staging;
    a.results = (garg_ray_probe_result*)((char*)staging + results_offset);
    a.count = count;
    HIP_TRY(hipMemcpyAsync((void*)a.probes, probes, (size_t)count * sizeof(garg_ray_probe), hipMemcpyHostToDevice, s),
            "uploading probes");
    HIP_TRY(launch_probe(a, integrator, s), "launching probeRays");
    HIP_TRY(hipMemcpyAsync(results, a.results, (size_t)count * sizeof(garg_ray_probe_result), hipMemcpyDeviceToHost, s),
            "reading probe results");
    return sync_stream(s, "waiting for the probes");
}

}  // namespace

int garg_probe(garg_ctx* ctx, garg_stream* stream, const garg_ray_probe* probes, uint32_t count,
               const garg_march_uniforms* uniforms, garg_tables* tables, uint32_t integrator,
               garg_ray_probe_result* results) {
    int rc = check_probes(ctx, stream, probes, count, uniforms, tables, integrator, results, "garg_probe");
    if (rc) return rc;
    uint64_t offset = align256((uint64_t)count * sizeof(garg_ray_probe));
    uint64_t need = offset + (uint64_t)count * sizeof(garg_ray_probe_result);
    if (need > ctx->probe_scratch_bytes) {
        ctx_free(ctx, ctx->probe_scratch, ctx->probe_scratch_bytes);
        ctx->probe_scratch = nullptr;
        ctx->probe_scratch_bytes = 0;
        rc = ctx_alloc(ctx, need, &ctx->probe_scratch, "the probe buffers");
        if (rc) return rc;
        ctx->probe_scratch_bytes = need;
    }
    return run_probes(ctx, stream, probes, count, uniforms, tables, integrator, results, ctx->probe_scratch, offset);
}

int garg_probe_noalloc(garg_ctx* ctx, garg_stream* stream, const garg_ray_probe* probes, uint32_t count,
                       const garg_march_uniforms* uniforms, garg_tables* tables, uint32_t integrator,
                       garg_ray_probe_result* results) {
    int rc = check_probes(ctx, stream, probes, count, uniforms, tables, integrator, results, "garg_probe_noalloc");
    if (rc) return rc;
    REQUIRE(count <= GARG_PROBE_NOALLOC_MAX,
            "garg_probe_noalloc takes at most %u probes per call (its staging is a fixed array), not %u.",
            GARG_PROBE_NOALLOC_MAX, count);
    void* staging = ctx->device < MAX_DEVICES ? g_scratch_ptr[ctx->device] : nullptr;
    if (!staging)
        return fail(GARG_E_STATE, "garg_probe_noalloc has no staging array on device %d.", ctx->device);
    return run_probes(ctx, stream, probes, count, uniforms, tables, integrator, results, staging, NOALLOC_PROBES_BYTES);
}

uint32_t garg_request_rays(uint32_t width, uint32_t height) { return ((width + 7u) / 8u) * ((height + 7u) / 8u) * 64u; }

int garg_batch_config_default(garg_ctx* ctx, garg_batch_config* out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    // Whole-tile claims at five waves per SIMD: the fastest policy on frames
    // whose neighbouring rays take similar step counts (lane refill pays
    // only on strongly skewed mixes; the autotuner picks per workload).
    out->block = 64;
    out->waves_per_simd = 5;
    out->compaction_period = 32;
    out->refill_threshold = 0;
    out->kerr_schild = 0;
    return GARG_OK;
}

// Bands a request is split into in the batch queue (at most, and only while
// the table has room for every band).
static const uint32_t BATCH_BANDS = 4;
// Largest Schwarzschild step budget marched without tile aging.
static const uint32_t SCHWARZSCHILD_SHORT_BUDGET = 256;

int garg_march_batch(garg_ctx* ctx, garg_stream* stream, const garg_march_request* requests, uint32_t count,
                     const garg_batch_config* config, garg_march_counters* total) {
    REQUIRE(ctx && requests && config, "ctx, requests and config are required.");
    REFUSE_WHILE_CAPTURING(ctx, stream, "garg_march_batch");
    REQUIRE(count > 0 && count <= GARG_BATCH_MAX_REQUESTS, "a batch holds 1 to %d requests, not %u.",
            GARG_BATCH_MAX_REQUESTS, count);
    REQUIRE(config->block == 64 || config->block == 128 || config->block == 256, "block must be 64, 128 or 256.");
    REQUIRE(config->compaction_period > 0, "the compaction period must be positive.");
    REQUIRE(config->refill_threshold <= 64, "the refill threshold is a lane count of at most 64.");
    REQUIRE(config->kerr_schild <= 2u, "kerr_schild is 0, 1 or 2, not %u.", config->kerr_schild);
    for (uint32_t i = 0; i < count; ++i) {
        const garg_march_request& r = requests[i];
        int rc = check_uniforms(&r.uniforms);
        if (rc) return fail(rc, "request %u: %s", i, g_error.c_str());
        REQUIRE(r.targets, "request %u has no targets.", i);
        REQUIRE(r.targets->width == r.uniforms.resolution[0] && r.targets->height == r.uniforms.resolution[1],
                "request %u: the targets do not match its resolution.", i);
        REQUIRE(r.uniforms.spin > 0.0f || r.uniforms.sphere_radius <= 0.0f || r.tables,
                "request %u enables the sphere jump but has no tables.", i);
        REQUIRE(r.ray_begin <= garg_request_rays(r.uniforms.resolution[0], r.uniforms.resolution[1]),
                "request %u starts past its last ray.", i);
    }
    hipStream_t s = stream_of(ctx, stream);
    // The device tables, queues and totals are reused: a call on another
    // stream first waits for the previous call's launches.
    HIP_TRY(hipStreamWaitEvent(s, ctx->batch_done, 0), "ordering after the previous batch");
    HIP_TRY(launch_fill_zero(ctx->batch_totals, sizeof(garg_march_counters), s), "clearing batch counters");
    uint32_t waves = config->waves_per_simd ? config->waves_per_simd : 4u;
    uint32_t grid = std::max(1u, ctx->info.compute_units * 4u * waves / (config->block / 64u));
    for (int part = 0; part < 2; ++part) {
        const bool kerr = part == 0;
        // This half of the pinned staging is reused once its previous copy ran.
        HIP_TRY(hipEventSynchronize(ctx->batch_staging_free[part]), "waiting for the staging buffer");
        BatchRequest* staging = ctx->batch_staging + (size_t)part * GARG_BATCH_MAX_REQUESTS;
        BatchRequest* table = ctx->batch_requests + (size_t)part * GARG_BATCH_MAX_REQUESTS;
        // This part's requests and their ray ranges.
        uint32_t members[GARG_BATCH_MAX_REQUESTS], member_rays[GARG_BATCH_MAX_REQUESTS];
        uint32_t m = 0, max_cap = 0;
        for (uint32_t i = 0; i < count; ++i) {
            const garg_march_request& r = requests[i];
            if ((r.uniforms.spin > 0.0f) != kerr) continue;
            uint32_t all = garg_request_rays(r.uniforms.resolution[0], r.uniforms.resolution[1]);
            uint32_t rays = all - r.ray_begin;
            if (r.ray_count) rays = std::min(rays, r.ray_count);
            if (rays == 0) continue;
            members[m] = i;
            member_rays[m++] = rays;
            max_cap = std::max(max_cap, r.uniforms.geodesic.step_cap);
        }
        // With whole-tile claims, a lone request of a whole image at most two
        // and a half rounds of waves deep goes to the one-launch-per-frame
        // kernel instead: its one-wave workgroups fit more waves per SIMD
        // than the persistent tile loop, which pays only over more rounds
        // (or more requests). The rays take the same steps and give the same
        // bits either way. A lane-refill configuration always runs the
        // persistent kernel.
        if (m == 1u && config->refill_threshold == 0u) {
            const garg_march_request& r = requests[members[0]];
            const uint32_t integrator = !kerr                       ? GARG_INTEGRATOR_FAST
                                        : config->kerr_schild == 1u ? GARG_INTEGRATOR_KERR_SCHILD
                                        : config->kerr_schild == 2u ? GARG_INTEGRATOR_POLAR
                                                                    : GARG_INTEGRATOR_FAST;
            if (!ctx->image_resident[integrator]) ctx->image_resident[integrator] = march_image_resident_blocks(integrator);
            const uint64_t slots = (uint64_t)ctx->image_resident[integrator] * ctx->info.compute_units;
            const uint64_t tiles = garg_request_rays(r.uniforms.resolution[0], r.uniforms.resolution[1]) / 64u;
            if (r.ray_begin == 0u &&
                member_rays[0] == garg_request_rays(r.uniforms.resolution[0], r.uniforms.resolution[1]) &&
                r.uniforms.tile_origin[0] == 0u && r.uniforms.tile_origin[1] == 0u && slots > 0u &&
                2u * tiles <= 5u * slots) {
                MarchImageArgs a;
                a.uniforms = device_uniforms(&r.uniforms, r.volume);
                a.scene = make_scene(ctx, r.volume, kerr ? nullptr : r.tables);
                a.color = r.targets->color;
                a.debug = r.targets->debug;
                a.depth = r.targets->depth;
                a.counters = ctx->lone_counters;
                a.extent_w = r.uniforms.resolution[0];
                a.extent_h = r.uniforms.resolution[1];
                HIP_TRY(launch_fill_zero(ctx->lone_counters, sizeof(garg_march_counters), s), "clearing counters");
                HIP_TRY(launch_march_image(a, integrator, s), "launching marchImage for a lone request");
                HIP_TRY(launch_add_counters(r.targets->counters, ctx->batch_totals, ctx->lone_counters, s),
                        "adding the request's counters");
                continue;
            }
        }
        // Each request is queued as up to BATCH_BANDS bands of whole tiles
        // (horizontal bands of its image), the middle bands first and the
        // requests' bands interleaved: the expensive tiles near a shadow or
        // photon ring start early instead of at the end of the queue, so the
        // launch does not end on one request's slowest tiles. The queue
        // order changes no ray's result.
        const uint32_t bands = std::max(1u, std::min(BATCH_BANDS, m ? GARG_BATCH_MAX_REQUESTS / m : 1u));
        uint32_t band_order[BATCH_BANDS];
        for (uint32_t k = 0; k < bands; ++k) band_order[k] = k;
        std::stable_sort(band_order, band_order + bands, [&](uint32_t x, uint32_t y) {
            return std::fabs((double)x - 0.5 * (bands - 1)) < std::fabs((double)y - 0.5 * (bands - 1));
        });
        uint32_t n = 0;
        uint64_t queue = 0;
        for (uint32_t rank = 0; rank < bands; ++rank) {
            for (uint32_t j = 0; j < m; ++j) {
                const garg_march_request& r = requests[members[j]];
                const uint32_t rays = member_rays[j];
                const uint32_t step = ((rays + bands - 1) / bands + 63u) & ~63u;
                const uint32_t first = band_order[rank] * step;
                if (first >= rays) continue;
                BatchRequest& b = staging[n++];
                b.uniforms = device_uniforms(&r.uniforms, r.volume);
                b.scene = make_scene(ctx, r.volume, kerr ? nullptr : r.tables);
                b.color = r.targets->color;
                b.debug = r.targets->debug;
                b.depth = r.targets->depth;
                b.counters = r.targets->counters;
                b.tiles_x = (r.uniforms.resolution[0] + 7u) / 8u;
                b.ray_begin = r.ray_begin + first;
                b.queue_begin = queue;
                b.ray_count = std::min(step, rays - first);
                b.pad = 0;
                // 64-aligned ranges: a claim of 64 rays never spans two entries.
                queue += ((uint64_t)b.ray_count + 63u) & ~(uint64_t)63u;
            }
        }
        if (n == 0) continue;
        HIP_TRY(hipMemcpyAsync(table, staging, sizeof(BatchRequest) * n, hipMemcpyHostToDevice, s), "uploading the batch");
        HIP_TRY(hipEventRecord(ctx->batch_staging_free[part], s), "recording an event");
        BatchArgs a;
        a.requests = table;
        a.request_count = n;
        a.kind = !kerr                       ? BATCH_SCHWARZSCHILD
                 : config->kerr_schild == 1u ? BATCH_KERR_SCHILD
                 : config->kerr_schild == 2u ? BATCH_KERR_POLAR
                                             : BATCH_KERR;
        a.total_rays = queue;
        a.queue = ctx->queue + (size_t)part * 8 * 16;
        a.totals = ctx->batch_totals;
        a.compaction_period = config->compaction_period;
        a.refill_threshold = config->refill_threshold;
        // Tile aging keeps long-lived persistent waves from starving (the
        // SIMD's issue arbitration favours the same wave slots); waves of
        // Schwarzschild batches within the interactive 256-step budget live
        // too briefly for that, and without it the kernel fits five waves
        // per SIMD.
        a.aging = kerr || max_cap > SCHWARZSCHILD_SHORT_BUDGET ? 1u : 0u;
        uint32_t resident =
            march_batch_resident_blocks(a.kind, a.aging != 0u, config->refill_threshold == 0u, config->block);
        uint32_t launch_grid = resident ? std::min(grid, resident * ctx->info.compute_units) : grid;
        HIP_TRY(launch_march_batch(a, launch_grid, config->block, s), "launching the batched march");
    }
    HIP_TRY(hipEventRecord(ctx->batch_done, s), "recording an event");
    if (total) {
        HIP_TRY(hipMemcpyAsync(total, ctx->batch_totals, sizeof(garg_march_counters), hipMemcpyDeviceToHost, s),
                "reading batch counters");
        return sync_stream(s, "waiting for the batch");
    }
    return GARG_OK;
}

// ---- bake / walk ---------------------------------------------------------------

uint64_t garg_bake_bytes(uint32_t width, uint32_t height, uint32_t capacity) {
    uint64_t rays = (uint64_t)width * height;
    return align256(rays * capacity * 8u) + align256(rays * 16u) + 256u;
}

int garg_bake_create(garg_ctx* ctx, uint32_t width, uint32_t height, uint32_t capacity, garg_bake** out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    REQUIRE(width > 0 && height > 0 && capacity > 0 && capacity <= 4096, "bad bake size.");
    garg_bake* b = new garg_bake();
    b->ctx = ctx;
    b->width = width;
    b->height = height;
    b->capacity = capacity;
    b->bytes = garg_bake_bytes(width, height, capacity);
    int rc = ctx_alloc(ctx, b->bytes, &b->base, "the bake buffers");
    if (rc) {
        delete b;
        return rc;
    }
    uint64_t rays = (uint64_t)width * height;
    char* p = static_cast<char*>(b->base);
    b->samples = (h4*)p;
    b->headers = (uint4*)(p + align256(rays * capacity * 8u));
    b->overflow = (uint32_t*)(p + align256(rays * capacity * 8u) + align256(rays * 16u));
    hipError_t e = launch_fill_zero(b->overflow, 256, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    HIP_TRY(e, "clearing the bake counter");
    *out = b;
    return GARG_OK;
}

int garg_bake_destroy(garg_bake* b) {
    if (!b) return GARG_OK;
    REFUSE_WHILE_CAPTURING(b->ctx, (garg_stream*)nullptr, "garg_bake_destroy");
    (void)hipStreamSynchronize(b->ctx->stream);
    ctx_free(b->ctx, b->base, b->bytes);
    delete b;
    return GARG_OK;
}

int garg_bake_rows(garg_bake* b, garg_stream* stream, const garg_march_uniforms* u, garg_tables* tables, uint32_t row0,
                   uint32_t rows) {
    REQUIRE(b, "bake is required.");
    int rc = check_uniforms(u);
    if (rc) return rc;
    REQUIRE(u->spin == 0.0f, "the bake strategy integrates Schwarzschild rays only; spin must be 0.");
    REQUIRE(u->resolution[0] == b->width && u->resolution[1] == b->height, "the bake buffers are %ux%u.", b->width,
            b->height);
    REQUIRE(row0 < b->height && rows > 0, "rows [%u, %u) lie outside the image.", row0, row0 + rows);
    REQUIRE(u->sphere_radius <= 0.0f || tables, "the uniforms enable the sphere jump but no tables were given.");
    hipStream_t s = stream_of(b->ctx, stream);
    if (row0 == 0) HIP_TRY(launch_fill_zero(b->overflow, 256, s), "clearing the bake counter");
    BakeArgs a;
    a.uniforms = device_uniforms(u, nullptr);
    a.scene = make_scene(b->ctx, nullptr, tables);
    a.samples = b->samples;
    a.headers = b->headers;
    a.overflow = b->overflow;
    a.row0 = row0;
    a.rows = std::min(rows, b->height - row0);
    a.capacity = b->capacity;
    HIP_TRY(launch_bake(a, s), "launching bakeRays");
    return GARG_OK;
}

int garg_bake_walk(garg_bake* b, garg_stream* stream, const garg_march_uniforms* u, garg_volume* volume,
                   garg_targets* targets) {
    REQUIRE(b && targets, "bake and targets are required.");
    int rc = check_uniforms(u);
    if (rc) return rc;
    REQUIRE(targets->width == b->width && targets->height == b->height, "the targets do not match the bake.");
    REQUIRE(u->resolution[0] == b->width && u->resolution[1] == b->height, "the uniforms do not match the bake.");
    WalkArgs a;
    a.uniforms = device_uniforms(u, volume);
    a.scene = make_scene(b->ctx, volume, nullptr);
    a.samples = b->samples;
    a.headers = b->headers;
    a.color = targets->color;
    a.depth = targets->depth;
    a.capacity = b->capacity;
    HIP_TRY(launch_walk(a, stream_of(b->ctx, stream)), "launching walkRays");
    return GARG_OK;
}

int garg_bake_overflow(garg_bake* b, uint32_t* rays) {
    REQUIRE(b && rays, "bake and rays are required.");
    REFUSE_WHILE_CAPTURING(b->ctx, (garg_stream*)nullptr, "garg_bake_overflow");
    HIP_TRY(hipDeviceSynchronize(), "waiting for the device");
    HIP_TRY(hipMemcpy(rays, b->overflow, 4, hipMemcpyDeviceToHost), "reading the overflow count");
    return GARG_OK;
}

int garg_bake_read(garg_bake* b, uint32_t* headers, uint16_t* samples) {
    REQUIRE(b, "bake is required.");
    REFUSE_WHILE_CAPTURING(b->ctx, (garg_stream*)nullptr, "garg_bake_read");
    HIP_TRY(hipDeviceSynchronize(), "waiting for the device");
    uint64_t rays = (uint64_t)b->width * b->height;
    if (headers) HIP_TRY(hipMemcpy(headers, b->headers, rays * 16u, hipMemcpyDeviceToHost), "reading bake headers");
    if (samples)
        HIP_TRY(hipMemcpy(samples, b->samples, rays * b->capacity * 8u, hipMemcpyDeviceToHost), "reading bake samples");
    return GARG_OK;
}

// ---- present -------------------------------------------------------------------

int garg_tonemap(garg_targets* t, garg_stream* stream, float exposure, uint32_t debug_view, float drift_budget,
                 uint8_t* rgba8) {
    REQUIRE(t && rgba8, "targets and rgba8 are required.");
    REFUSE_WHILE_CAPTURING(t->ctx, stream, "garg_tonemap");
    REQUIRE(drift_budget > 0.0f || !debug_view, "the debug view needs a positive drift budget.");
    hipStream_t s = stream_of(t->ctx, stream);
    TonemapArgs a{t->color, t->debug, t->rgba8, t->width * t->height, exposure, debug_view ? 1u : 0u, drift_budget};
    HIP_TRY(launch_tonemap(a, s), "launching the tone curve");
    HIP_TRY(hipMemcpyAsync(rgba8, t->rgba8, (size_t)t->width * t->height * 4u, hipMemcpyDeviceToHost, s),
            "reading the tone mapped image");
    return sync_stream(s, "waiting for the tone curve");
}

// ---- FP64 kernels ----------------------------------------------------------------

int garg_deflection_table(garg_ctx* ctx, garg_stream* stream, double r_obs, uint32_t count, double b_lo, double b_hi,
                          uint32_t intervals, double* b_out, double* alpha_out, double* sweep_out) {
    REQUIRE(ctx && b_out && alpha_out && sweep_out, "ctx and the three outputs are required.");
    REFUSE_WHILE_CAPTURING(ctx, stream, "garg_deflection_table");
    REQUIRE(count >= 2 && count <= (1u << 20), "count must be in [2, 2^20].");
    REQUIRE(b_lo > B_CRITICAL && b_hi > b_lo, "the impact parameters must lie above 3 sqrt(3) and increase.");
    REQUIRE(r_obs > 3.0 && intervals > 0, "r_obs must exceed 3 and intervals must be positive.");
    double* buf = nullptr;
    int rc = ctx_alloc(ctx, (uint64_t)count * 24u, (void**)&buf, "the deflection table");
    if (rc) return rc;
    hipStream_t s = stream_of(ctx, stream);
    hipError_t e = launch_far_field(buf, buf + count, buf + 2 * count, count, r_obs, b_lo, b_hi, intervals, s);
    if (e == hipSuccess) e = hipMemcpyAsync(b_out, buf, count * 8u, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(alpha_out, buf + count, count * 8u, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(sweep_out, buf + 2 * count, count * 8u, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ctx_free(ctx, buf, (uint64_t)count * 24u);
    HIP_TRY(e, "building the deflection table");
    return GARG_OK;
}

int garg_probe_fp64(garg_ctx* ctx, garg_stream* stream, const garg_ray_probe* probes, uint32_t count,
                    const garg_geodesic_settings* settings, garg_ray_probe_fp64_result* results) {
    REQUIRE(ctx && settings && (count == 0 || (probes && results)), "ctx, settings, probes and results are required.");
    REFUSE_WHILE_CAPTURING(ctx, stream, "garg_probe_fp64");
    REQUIRE(count <= 65536u, "at most 65536 FP64 probes per call.");
    const garg_geodesic_settings& g = *settings;
    REQUIRE(g.step_min > 0.0f && g.step_max >= g.step_min && g.empty_step_max >= g.step_min && g.step_factor > 0.0f &&
                g.empty_step_factor > 0.0f,
            "the step policy needs positive factors and step_min <= step_max.");
    REQUIRE(g.capture_radius > R_SCHWARZSCHILD && g.escape_radius > g.capture_radius,
            "the capture radius must lie outside r = 2 and inside the escape radius.");
    REQUIRE(g.step_cap > 0u && g.step_cap <= (1u << 24), "the step cap must be in [1, 2^24].");
    for (uint32_t i = 0; i < count; ++i) {
        const float* o = probes[i].origin;
        const double r = std::sqrt((double)o[0] * o[0] + (double)o[1] * o[1] + (double)o[2] * o[2]);
        const float* d = probes[i].direction;
        REQUIRE(r > g.capture_radius && std::isfinite(r), "probe %u starts at r = %g, inside the capture radius.", i, r);
        REQUIRE(d[0] != 0.0f || d[1] != 0.0f || d[2] != 0.0f, "probe %u has no direction.", i);
    }
    if (count == 0) return GARG_OK;
    const ProbeSettings64 g64{g.step_factor, g.step_min, g.step_max, g.capture_radius, g.escape_radius, g.step_cap,
                              g.empty_step_factor, g.empty_step_max};
    const uint64_t in_bytes = align256((uint64_t)count * sizeof(garg_ray_probe));
    const uint64_t bytes = in_bytes + (uint64_t)count * sizeof(garg_ray_probe_fp64_result);
    void* buf = nullptr;
    int rc = ctx_alloc(ctx, bytes, &buf, "the FP64 probes");
    if (rc) return rc;
    garg_ray_probe* d_in = (garg_ray_probe*)buf;
    garg_ray_probe_fp64_result* d_out = (garg_ray_probe_fp64_result*)((char*)buf + in_bytes);
    hipStream_t s = stream_of(ctx, stream);
    hipError_t e = hipMemcpyAsync(d_in, probes, (size_t)count * sizeof(garg_ray_probe), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_probe_fp64(d_in, count, g64, d_out, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(results, d_out, (size_t)count * sizeof(garg_ray_probe_fp64_result), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ctx_free(ctx, buf, bytes);
    HIP_TRY(e, "integrating the FP64 probes");
    return GARG_OK;
}

int garg_render_fp64(garg_ctx* ctx, garg_stream* stream, const garg_volume* volume,
                     const garg_render_camera_fp64* camera, uint32_t width, uint32_t height,
                     const garg_geodesic_settings* settings, double opacity, uint32_t redshift, double* rgb) {
    REQUIRE(ctx && volume && camera && settings && rgb, "ctx, volume, camera, settings and rgb are required.");
    REQUIRE(volume->ctx == ctx, "the volume belongs to another context.");
    REFUSE_WHILE_CAPTURING(ctx, stream, "garg_render_fp64");
    REQUIRE(width > 0 && height > 0 && (uint64_t)width * height <= (1u << 22),
            "the FP64 render takes 1 to 2^22 pixels (%u x %u asked).", width, height);
    const garg_geodesic_settings& g = *settings;
    REQUIRE(g.step_min > 0.0f && g.step_max >= g.step_min && g.empty_step_max >= g.step_min && g.step_factor > 0.0f &&
                g.empty_step_factor > 0.0f,
            "the step policy needs positive factors and step_min <= step_max.");
    REQUIRE(g.capture_radius > R_SCHWARZSCHILD && g.escape_radius > g.capture_radius,
            "the capture radius must lie outside r = 2 and inside the escape radius.");
    REQUIRE(g.step_cap > 0u && g.step_cap <= (1u << 24), "the step cap must be in [1, 2^24].");
    REQUIRE(opacity >= 0.0 && std::isfinite(opacity), "the opacity must be finite and non-negative.");
    const double* p = camera->position;
    const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    REQUIRE(r > g.capture_radius && std::isfinite(r), "the camera sits at r = %g, inside the capture radius.", r);
    Render64 a{};
    a.voxels = volume->voxels;
    a.size = volume->size;
    a.half_xy = volume->half_xy;
    a.half_z = volume->half_z;
    for (int k = 0; k < 3; ++k) {
        a.position[k] = camera->position[k];
        a.right[k] = camera->right[k];
        a.up[k] = camera->up[k];
        a.forward[k] = camera->forward[k];
    }
    a.tan_half_fov[0] = camera->tan_half_fov[0];
    a.tan_half_fov[1] = camera->tan_half_fov[1];
    a.width = width;
    a.height = height;
    a.opacity = opacity;
    a.redshift = redshift ? 1u : 0u;
    a.settings = ProbeSettings64{g.step_factor, g.step_min, g.step_max, g.capture_radius, g.escape_radius, g.step_cap,
                                 g.empty_step_factor, g.empty_step_max};
    const uint64_t bytes = (uint64_t)width * height * 3u * sizeof(double);
    void* buf = nullptr;
    int rc = ctx_alloc(ctx, bytes, &buf, "the FP64 render");
    if (rc) return rc;
    a.rgb = (double*)buf;
    hipStream_t s = stream_of(ctx, stream);
    hipError_t e = launch_render_fp64(a, s);
    if (e == hipSuccess) e = hipMemcpyAsync(rgb, buf, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ctx_free(ctx, buf, bytes);
    HIP_TRY(e, "rendering in FP64");
    return GARG_OK;
}

int garg_shadow_boundary(garg_ctx* ctx, garg_stream* stream, double spin, double r_obs, double inclination_deg,
                         uint32_t azimuths, uint32_t rounds, uint32_t rays_per_round, uint32_t step_cap,
                         double* edge_out, double* xi_out) {
    REQUIRE(ctx && edge_out, "ctx and edge_out are required.");
    REFUSE_WHILE_CAPTURING(ctx, stream, "garg_shadow_boundary");
    REQUIRE(spin >= 0.0 && spin < 1.0, "spin must be in [0, 1).");
    REQUIRE(r_obs > 5.0, "the observer must sit outside r = 5.");
    REQUIRE(inclination_deg > 0.0 && inclination_deg < 180.0, "the inclination must be in (0, 180) degrees.");
    REQUIRE(azimuths > 0 && azimuths <= 65536 && rounds > 0 && rounds <= 16, "bad azimuth or round count.");
    REQUIRE(rays_per_round == 64, "the multi-section runs one wave64 per azimuth: 64 rays per round.");
    REQUIRE(step_cap > 0 && step_cap <= (1u << 20), "bad step cap.");
    double* buf = nullptr;
    int rc = ctx_alloc(ctx, (uint64_t)azimuths * 16u, (void**)&buf, "the shadow boundary");
    if (rc) return rc;
    hipStream_t s = stream_of(ctx, stream);
    hipError_t e = launch_shadow(buf, buf + azimuths, azimuths, spin, r_obs, inclination_deg * M_PI / 180.0, rounds,
                                 rays_per_round, step_cap, s);
    if (e == hipSuccess) e = hipMemcpyAsync(edge_out, buf, azimuths * 8u, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && xi_out) e = hipMemcpyAsync(xi_out, buf + azimuths, azimuths * 8u, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ctx_free(ctx, buf, (uint64_t)azimuths * 16u);
    HIP_TRY(e, "bisecting the shadow boundary");
    return GARG_OK;
}

}  // extern "C"

// ---- graphs ------------------------------------------------------------------

struct garg_graph {
    garg_ctx* ctx;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

int garg_graph_begin(garg_ctx* ctx, garg_stream* stream) {
    REQUIRE(ctx, "ctx is required.");
    hipStream_t s = stream_of(ctx, stream);
    if (capturing(s)) return fail(GARG_E_STATE, "the stream is already capturing a graph.");
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "beginning a graph capture");
    return GARG_OK;
}

int garg_graph_end(garg_ctx* ctx, garg_stream* stream, garg_graph** out) {
    REQUIRE(ctx && out, "ctx and out are required.");
    hipStream_t s = stream_of(ctx, stream);
    if (!capturing(s)) return fail(GARG_E_STATE, "the stream is not capturing a graph.");
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture(s, &g), "ending the graph capture");
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(GARG_E_DEVICE, "instantiating the graph: %s", hipGetErrorString(e));
    }
    *out = new garg_graph{ctx, g, exec};
    return GARG_OK;
}

int garg_graph_launch(garg_graph* graph, garg_stream* stream) {
    REQUIRE(graph, "graph is required.");
    HIP_TRY(hipGraphLaunch(graph->exec, stream_of(graph->ctx, stream)), "launching the graph");
    return GARG_OK;
}

int garg_graph_destroy(garg_graph* graph) {
    if (!graph) return GARG_OK;
    (void)hipGraphExecDestroy(graph->exec);
    (void)hipGraphDestroy(graph->graph);
    delete graph;
    return GARG_OK;
}

int garg_particles_dispatches(garg_particles* p, uint32_t* out) {
    REQUIRE(p && out, "particles and out are required.");
    REFUSE_WHILE_CAPTURING(p->ctx, (garg_stream*)nullptr, "garg_particles_dispatches");
    HIP_TRY(hipDeviceSynchronize(), "waiting for the device");
    HIP_TRY(hipMemcpy(out, p->view.dispatches, 4, hipMemcpyDeviceToHost), "reading the dispatch count");
    return GARG_OK;
}

// ---- measurement ---------------------------------------------------------------

int garg_atomic_rate(garg_ctx* ctx, garg_stream* stream, uint32_t words, uint32_t adds_per_thread, uint32_t threads,
                     double* adds_per_second) {
    REQUIRE(ctx && adds_per_second, "ctx and adds_per_second are required.");
    REQUIRE(words > 0 && adds_per_thread > 0 && threads > 0, "words, adds and threads must be positive.");
    REFUSE_WHILE_CAPTURING(ctx, stream, "garg_atomic_rate");
    hipStream_t s = stream_of(ctx, stream);
    void* buffer = nullptr;
    int rc = ctx_alloc(ctx, (uint64_t)words * 4u, &buffer, "the atomic buffer");
    if (rc) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = launch_fill_zero(buffer, (uint64_t)words * 4u, s);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = launch_atomic_rate((uint32_t*)buffer, words, adds_per_thread, threads, s);  // warm
    if (e == hipSuccess) e = hipEventRecord(a, s);
    if (e == hipSuccess) e = launch_atomic_rate((uint32_t*)buffer, words, adds_per_thread, threads, s);
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    ctx_free(ctx, buffer, (uint64_t)words * 4u);
    HIP_TRY(e, "timing the atomic kernel");
    *adds_per_second = ms > 0.0f ? (double)threads * adds_per_thread / (ms * 1e-3) : 0.0;
    return GARG_OK;
}

int garg_valu_rate(garg_ctx* ctx, garg_stream* stream, uint32_t form, uint32_t iterations, uint32_t threads,
                   double* flops_per_second) {
    REQUIRE(ctx && flops_per_second, "ctx and flops_per_second are required.");
    REQUIRE(form <= GARG_VALU_FMA_SGPR, "unknown FMA operand form %u.", form);
    REQUIRE(iterations > 0 && threads > 0, "iterations and threads must be positive.");
    REFUSE_WHILE_CAPTURING(ctx, stream, "garg_valu_rate");
    hipStream_t s = stream_of(ctx, stream);
    void* buffer = nullptr;
    uint64_t bytes = align256((uint64_t)threads * 4u);
    int rc = ctx_alloc(ctx, bytes, &buffer, "the FMA output");
    if (rc) return rc;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = launch_valu_rate((float*)buffer, form, threads, iterations, s);  // warm
    if (e == hipSuccess) e = hipEventRecord(a, s);
    if (e == hipSuccess) e = launch_valu_rate((float*)buffer, form, threads, iterations, s);
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    ctx_free(ctx, buffer, bytes);
    HIP_TRY(e, "timing the FMA kernel");
    *flops_per_second = ms > 0.0f ? 2.0 * 32.0 * threads * (double)iterations / (ms * 1e-3) : 0.0;
    return GARG_OK;
}

