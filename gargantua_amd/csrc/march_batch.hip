// The batched persistent march for gfx950. Rays of many render requests,
// each in 8 x 8 tile order, form one queue split into eight contiguous
// ranges, one per XCD (workgroups are dispatched round-robin over the eight
// XCDs, so blockIdx % 8 names the XCD and its L2). A wave keeps 64 rays in
// registers, marches them a compaction period at a time, and when at most
// `refill_threshold` lanes are still live it refills the finished lanes with
// the next rays of its XCD's range in one atomic claim, stealing from the
// other ranges once its own is empty (Aila and Laine 2009, persistent
// threads with dynamic fetch). Every ray is marched by the same per-ray
// state machine as the one-launch-per-frame kernel, so its result does not
// depend on which lane, wave or batch carries it.
#include "device_march.hpp"
#include "launch.hpp"

namespace garg {

constexpr int XCDS = 8;
// Real-time age (100 MHz ticks, 200 us) at which a tile first rises in issue
// priority; it rises again at twice and four times that age.
constexpr uint64_t TILE_AGE_STEP = 20000;
constexpr int QUEUE_STRIDE = 16;  // 128 bytes between queue heads

GARG_DEV uint64_t wave_broadcast(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Request holding global ray g (requests sorted by queue_begin).
GARG_DEV uint32_t find_request(const BatchArgs& a, uint64_t g) {
    uint32_t lo = 0, hi = a.request_count;
    while (hi - lo > 1u) {
        uint32_t mid = (lo + hi) >> 1;
        if (a.requests[mid].queue_begin <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct LaneRay {
    uint32_t request;
    uint32_t pixel;  // py * width + px
};

// Tile-order ray of a request to its pixel; false for the padding of the
// request's queue range and for edge-tile rays outside the image.
GARG_DEV bool ray_pixel(const BatchRequest& r, uint64_t g, uint32_t& px, uint32_t& py) {
    const uint32_t offset = (uint32_t)(g - r.queue_begin);
    if (offset >= r.ray_count) return false;
    uint32_t local = offset + r.ray_begin;
    uint32_t tile = local >> 6, lane = local & 63u;
    px = (tile % r.tiles_x) * 8u + (lane & 7u);
    py = (tile / r.tiles_x) * 8u + (lane >> 3);
    return px < r.uniforms.resolution[0] && py < r.uniforms.resolution[1];
}

struct CounterTable {
    uint32_t* counts;  // [request][4]: rays, over budget, exhausted, captured
    unsigned long long* steps;
};

GARG_DEV void finish_ray(const BatchRequest& r, CounterTable& t, uint32_t request, uint32_t pixel, f3 color,
                         float drift, uint32_t steps, uint32_t outcome, float depth) {
    store_half4(&r.color[pixel], color.x, color.y, color.z, 1.0f);
    store_half4(&r.debug[pixel], drift, (float)steps, (float)outcome, 1.0f);
    if (r.depth) r.depth[pixel] = depth;
    uint32_t* c = t.counts + 4u * request;
    atomicAdd(&c[0], 1u);
    if (drift > r.uniforms.drift_budget) atomicAdd(&c[1], 1u);
    if (outcome == GARG_RAY_EXHAUSTED) atomicAdd(&c[2], 1u);
    if (outcome == GARG_RAY_CAPTURED) atomicAdd(&c[3], 1u);
    if (steps) atomicAdd(&t.steps[request], (unsigned long long)steps);
}

// Raises the wave's issue priority as its tile ages: one level at
// TILE_AGE_STEP, two at twice that, three at four times (checked every 8
// steps from the 100 MHz real-time counter).
struct TileAging {
    uint64_t born;
    uint32_t priority = 0;
    GARG_DEV void operator()(uint32_t n) {
        if ((n & 7u) != 7u || priority >= 3u) return;
        const uint64_t age = __builtin_amdgcn_s_memrealtime() - born;
        if (age <= (TILE_AGE_STEP << priority)) return;
        ++priority;
        if (priority == 1u) __builtin_amdgcn_s_setprio(1);
        else if (priority == 2u) __builtin_amdgcn_s_setprio(2);
        else __builtin_amdgcn_s_setprio(3);
    }
};

// Sets the wave's issue priority from the age of its oldest live ray (the
// lane refill's counterpart of TileAging): 0 under TILE_AGE_STEP, then one
// level at that age, two at twice and three at four times. `started` is the
// low word of the 100 MHz real-time counter when the lane's ray started.
GARG_DEV void age_priority(bool live, uint32_t started) {
    const uint32_t now = (uint32_t)__builtin_amdgcn_s_memrealtime();
    uint32_t age = live ? now - started : 0u;
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) age = max(age, (uint32_t)__shfl_xor((int)age, offset));
    const uint32_t oldest = __builtin_amdgcn_readfirstlane(age);
    if (oldest > 4u * (uint32_t)TILE_AGE_STEP) __builtin_amdgcn_s_setprio(3);
    else if (oldest > 2u * (uint32_t)TILE_AGE_STEP) __builtin_amdgcn_s_setprio(2);
    else if (oldest > (uint32_t)TILE_AGE_STEP) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// The ray state machine a lane runs: Schwarzschild (with the jump), Kerr in
// Boyer-Lindquist coordinates (the FAST form) or Kerr in Kerr-Schild
// coordinates.
template <int KIND>
struct Lane;

template <class March>
struct KerrLane {
    March m;
    GARG_DEV bool start(const BatchRequest& r, uint32_t px, uint32_t py, CounterTable&, uint32_t, uint32_t) {
        const garg_march_uniforms& u = r.uniforms;
        f3 dir = camera_direction(u, px, py);
        m.begin(u.spin, kerr_launch(u.spin, load3(u.camera_position), dir));
        return true;
    }
    template <bool UNIFORM, class Tick = NoTick>
    GARG_DEV void run(const BatchRequest& r, uint32_t budget, Tick tick = Tick{}) {
        m.template run_loop<UNIFORM>(kerr_loop(r.uniforms, r.scene, true, r.uniforms.compensated != 0u), budget, tick);
    }
    GARG_DEV bool done() const { return m.done; }
    GARG_DEV void finish(const BatchRequest& r, CounterTable& t, uint32_t request, uint32_t pixel) {
        f3 color = finish_color(m.g, m.falls_in(), m.final_direction(), r.uniforms, r.scene);
        finish_ray(r, t, request, pixel, color, m.drift, m.steps, m.outcome,
                   fminf(m.g.depth() / (2.0f * GARG_F(R_ESCAPE)), 1.0f));
    }
};

template <>
struct Lane<BATCH_KERR> : KerrLane<KerrFastMarch> {};

template <>
struct Lane<BATCH_KERR_SCHILD> : KerrLane<KerrSchildMarch> {};

template <>
struct Lane<BATCH_KERR_POLAR> {
    KerrHybridMarch m;
    GARG_DEV bool start(const BatchRequest& r, uint32_t px, uint32_t py, CounterTable&, uint32_t, uint32_t) {
        const garg_march_uniforms& u = r.uniforms;
        f3 dir = camera_direction(u, px, py);
        m.begin(u.spin, kerr_launch(u.spin, load3(u.camera_position), dir));
        return true;
    }
    template <bool UNIFORM, class Tick = NoTick>
    GARG_DEV void run(const BatchRequest& r, uint32_t budget, Tick tick = Tick{}) {
        m.template run_loop<UNIFORM>(kerr_loop(r.uniforms, r.scene, true, r.uniforms.compensated != 0u), budget, tick);
    }
    GARG_DEV bool done() const { return m.done(); }
    GARG_DEV void finish(const BatchRequest& r, CounterTable& t, uint32_t request, uint32_t pixel) {
        f3 color = finish_color(m.gather(), m.falls_in(), m.final_direction(), r.uniforms, r.scene);
        finish_ray(r, t, request, pixel, color, m.max_drift(), m.step_count(), m.result(),
                   fminf(m.gather().depth() / (2.0f * GARG_F(R_ESCAPE)), 1.0f));
    }
};

template <>
struct Lane<BATCH_SCHWARZSCHILD> {
    SchwMarchT<true> m;
    // Rays that never enter the integration sphere finish at the jump.
    GARG_DEV bool start(const BatchRequest& r, uint32_t px, uint32_t py, CounterTable& t, uint32_t request,
                        uint32_t pixel) {
        const garg_march_uniforms& u = r.uniforms;
        f3 dir = camera_direction(u, px, py);
        Prepared p = prepare_ray(load3(u.camera_position), dir, u, r.scene);
        if (p.sky_only) {
            f3 color = u.star_brightness > 0.0f ? starfield(p.sky, u.star_seed, r.scene.blackbody) * u.star_brightness
                                                : f3{0, 0, 0};
            finish_ray(r, t, request, pixel, color, 0.0f, 0u, GARG_RAY_ESCAPED, 1.0f);
            return false;
        }
        m.begin(p.ray);
        return true;
    }
    template <bool UNIFORM, class Tick = NoTick>
    GARG_DEV void run(const BatchRequest& r, uint32_t budget, Tick tick = Tick{}) {
        if constexpr (UNIFORM) {
            // The wave's request: scalar copies of its uniforms and scene.
            const garg_march_uniforms u = wave_uniform(r.uniforms);
            const Scene sc = wave_uniform(r.scene);
            m.run(u, sc, true, u.compensated != 0u, budget, tick);
        } else {
            m.run(r.uniforms, r.scene, true, r.uniforms.compensated != 0u, budget);
        }
    }
    GARG_DEV bool done() const { return m.done; }
    GARG_DEV void finish(const BatchRequest& r, CounterTable& t, uint32_t request, uint32_t pixel) {
        bool in = m.falls_in();
        f3 color = finish_color(m.g, in, in ? f3{0, 0, 0} : m.sky_direction(r.uniforms, r.scene), r.uniforms, r.scene);
        finish_ray(r, t, request, pixel, color, ray_drift(m.ray), m.steps, m.outcome,
                   fminf(m.g.depth() / (2.0f * GARG_F(R_ESCAPE)), 1.0f));
    }
};

template <int BLOCK>
GARG_DEV void clear_counters(uint32_t* s_counts, unsigned long long* s_steps, uint32_t requests) {
    for (uint32_t i = threadIdx.x; i < requests * 4u; i += BLOCK) s_counts[i] = 0u;
    for (uint32_t i = threadIdx.x; i < requests; i += BLOCK) s_steps[i] = 0ull;
    __syncthreads();
}

// One global atomic per non-zero counter per block, into the request's
// counters and the batch totals.
template <int BLOCK>
GARG_DEV void flush_counters(const BatchArgs& a, const uint32_t* s_counts, const unsigned long long* s_steps) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.request_count; i += BLOCK) {
        garg_march_counters* c = a.requests[i].counters;
        uint32_t* cc = reinterpret_cast<uint32_t*>(c);
        uint32_t* tc = reinterpret_cast<uint32_t*>(a.totals);
        for (int k = 0; k < 4; ++k) {
            uint32_t v = s_counts[4u * i + k];
            if (v) {
                atomicAdd(&cc[k], v);
                atomicAdd(&tc[k], v);
            }
        }
        if (s_steps[i]) {
            atomicAdd((unsigned long long*)&c->steps, s_steps[i]);
            atomicAdd((unsigned long long*)&a.totals->steps, s_steps[i]);
        }
    }
}

// Claim up to `want` consecutive rays for the calling wave from its XCD's
// range, moving on to the other ranges once that is empty. Returns the
// count claimed (0 once every range is empty); wave-uniform.
GARG_DEV uint64_t claim_rays(const BatchArgs& a, uint32_t home, uint32_t& queue, bool& drained, uint64_t want,
                             uint64_t& base) {
    const uint32_t lane = __lane_id();
    while (!drained) {
        uint64_t b = 0;
        if (lane == 0) b = atomicAdd(&a.queue[queue * QUEUE_STRIDE], (unsigned long long)want);
        b = wave_broadcast(b);
        uint64_t end = a.queue[queue * QUEUE_STRIDE + 1];
        if (b < end) {
            base = b;
            return min(want, end - b);
        }
        queue = (queue + 1u) % XCDS;
        if (queue == home) drained = true;
    }
    return 0;
}

// Start the lane's ray g of the batch; false if it has nothing to march.
template <int KIND>
GARG_DEV bool start_ray(const BatchArgs& a, Lane<KIND>& ray, LaneRay& id, CounterTable& table, uint64_t g) {
    uint32_t r = find_request(a, g);
    const BatchRequest& req = a.requests[r];
    uint32_t px, py;
    if (!ray_pixel(req, g, px, py)) return false;
    id = LaneRay{r, py * req.uniforms.resolution[0] + px};
    return ray.start(req, px, py, table, r, id.pixel);
}

// March the live lanes `budget` steps, one request at a time (a wave almost
// always holds one): the request index is then wave-uniform, so its uniforms
// and scene are scalar loads held in SGPRs, and one copy of the loop serves
// every mix of requests. Finished rays are stored; returns whether the lane
// is still live.
template <int KIND>
GARG_DEV bool march_lanes(const BatchArgs& a, Lane<KIND>& ray, const LaneRay& id, CounterTable& table, bool live,
                          uint32_t budget) {
    bool pending = live;
    for (;;) {
        const uint64_t pending_mask = __ballot(pending);
        if (pending_mask == 0ull) break;
        const uint32_t lead = (uint32_t)__builtin_ctzll(pending_mask);
        const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)id.request, (int)lead);
        if (pending && id.request == r0) {
            const BatchRequest& req = a.requests[r0];
            ray.template run<true>(req, budget);
            if (ray.done()) {
                ray.finish(req, table, r0, id.pixel);
                live = false;
            }
            pending = false;
        }
    }
    return live;
}

// Lane-granular refill: a wave keeps marching while more than
// `refill_threshold` lanes are live, then tops its finished lanes up.
// Kerr-Schild rays carry more state than 128 registers hold next to the
// refill bookkeeping, so that kind runs three waves per SIMD here.
template <int KIND, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((
    amdgpu_waves_per_eu(KIND == BATCH_KERR_SCHILD || KIND == BATCH_KERR_POLAR ? 3 : 4))) void
march_batch_kernel(BatchArgs a) {
    __shared__ uint32_t s_counts[GARG_BATCH_MAX_REQUESTS * 4];
    __shared__ unsigned long long s_steps[GARG_BATCH_MAX_REQUESTS];
    clear_counters<BLOCK>(s_counts, s_steps, a.request_count);
    CounterTable table{s_counts, s_steps};

    const uint64_t below = (1ull << __lane_id()) - 1ull;
    const uint32_t home = blockIdx.x % XCDS;
    uint32_t queue = home;
    bool drained = false;

    Lane<KIND> ray;
    bool live = false;
    LaneRay id{0u, 0u};
    uint32_t started = 0;

    for (;;) {
        uint64_t dead = __ballot(!live);
        uint32_t live_count = 64u - (uint32_t)__popcll(dead);
        if (!drained && dead != 0ull && live_count <= a.refill_threshold) {
            // One ray per dead lane, in lane order.
            uint64_t base = 0;
            uint64_t got = claim_rays(a, home, queue, drained, (uint64_t)__popcll(dead), base);
            uint64_t rank = (uint64_t)__popcll(dead & below);
            if (!live && rank < got) {
                live = start_ray(a, ray, id, table, base + rank);
                started = (uint32_t)__builtin_amdgcn_s_memrealtime();
            }
        }
        if (__ballot(live) == 0ull) {
            if (drained) break;
            continue;
        }
        // Issue arbitration favours the same hardware wave slots for as long
        // as the waves live; the oldest rays go first, as in the tile policy.
        age_priority(live, started);
        live = march_lanes(a, ray, id, table, live, a.compaction_period);
    }
    flush_counters<BLOCK>(a, s_counts, s_steps);
}

// Tile-granular persistent march: a wave marches 64 consecutive rays (one
// 8 x 8 tile) to the end, then takes the next 64. The first round is
// static, so the launch does not open with every wave queueing on one
// device-scope atomic: wave k of workgroup b takes tile k * gridDim.x + b,
// so consecutive tiles go to consecutive workgroups (and so to different
// XCDs and CUs, as one-wave workgroups of a plain launch do) rather than
// the expensive neighbours of a ring or shadow edge all landing on one CU.
// Later rounds claim from a single head in ray order, which finishes the
// frame in the order the hardware would dispatch it and keeps the tail
// short. No ray state outlives a tile, so the kernel needs no more
// registers than the one-launch-per-frame march: five waves per SIMD for
// Boyer-Lindquist Kerr (96 VGPRs), four for Kerr-Schild. AGING = false drops
// the tile's priority aging: Schwarzschild batches of short budgets, whose
// waves live too briefly to starve, then fit five waves per SIMD, which the
// jump and gather do not next to the aging state.
template <int KIND, int BLOCK, bool AGING>
__global__ __launch_bounds__(BLOCK) __attribute__((
    amdgpu_waves_per_eu(KIND == BATCH_KERR || (KIND == BATCH_SCHWARZSCHILD && !AGING) ? 5 : 4))) void
march_tiles_kernel(BatchArgs a) {
    __shared__ uint32_t s_counts[GARG_BATCH_MAX_REQUESTS * 4];
    __shared__ unsigned long long s_steps[GARG_BATCH_MAX_REQUESTS];
    clear_counters<BLOCK>(s_counts, s_steps, a.request_count);
    CounterTable table{s_counts, s_steps};

    const uint32_t lane = __lane_id();
    uint64_t base = ((uint64_t)(threadIdx.x / 64u) * gridDim.x + blockIdx.x) * 64u;
    while (base < a.total_rays) {
        const uint64_t got = min((uint64_t)64, a.total_rays - base);
        // Request ranges are 64-aligned, so the whole tile is one request.
        const uint32_t r = find_request(a, base);
        const BatchRequest& req = a.requests[r];
        Lane<KIND> ray;
        uint32_t px = 0, py = 0;
        bool live = lane < got && ray_pixel(req, base + lane, px, py);
        const uint32_t pixel = py * req.uniforms.resolution[0] + px;
        if (live) live = ray.start(req, px, py, table, r, pixel);
        // Issue arbitration among the waves of a SIMD favours the same
        // hardware wave slots for as long as the waves live; a persistent
        // wave in a losing slot would keep one tile for the whole frame.
        // A tile therefore starts at the lowest priority and rises as it
        // ages in real time, so the oldest tiles finish first, as they do
        // when the hardware launches a fresh wave per tile.
        __builtin_amdgcn_s_setprio(0);
        if (live) {
            if constexpr (AGING) ray.template run<true>(req, 0xffffffffu, TileAging{__builtin_amdgcn_s_memrealtime()});
            else ray.template run<true>(req, 0xffffffffu);
        }
        if (live) ray.finish(req, table, r, pixel);
        // Claiming only now, not before the march, keeps the claim's result
        // out of the registers the march needs.
        uint64_t next = 0;
        if (lane == 0) next = atomicAdd(&a.queue[0], 64ull);
        base = wave_broadcast(next);
    }
    flush_counters<BLOCK>(a, s_counts, s_steps);
}

// Tile policy: one head, past the rays of the static first round.
__global__ void tile_queue_init_kernel(unsigned long long* queue, uint64_t first) {
    if (threadIdx.x == 0) queue[0] = first;
}

__global__ void queue_init_kernel(unsigned long long* queue, uint64_t total) {
    uint32_t q = threadIdx.x;
    if (q >= XCDS) return;
    // Contiguous, tile-aligned ranges.
    uint64_t tiles = (total + 63u) / 64u;
    uint64_t begin = min(total, (tiles * q / XCDS) * 64u);
    uint64_t end = min(total, (tiles * (q + 1u) / XCDS) * 64u);
    queue[q * QUEUE_STRIDE] = begin;
    queue[q * QUEUE_STRIDE + 1] = end;
}

template <int KIND, bool AGING>
static hipError_t launch_block(const BatchArgs& a, uint32_t grid, uint32_t block, hipStream_t stream) {
    const bool tiles = a.refill_threshold == 0u;
    switch (block) {
        case 64:
            if (tiles) march_tiles_kernel<KIND, 64, AGING><<<grid, 64, 0, stream>>>(a);
            else march_batch_kernel<KIND, 64><<<grid, 64, 0, stream>>>(a);
            break;
        case 128:
            if (tiles) march_tiles_kernel<KIND, 128, AGING><<<grid, 128, 0, stream>>>(a);
            else march_batch_kernel<KIND, 128><<<grid, 128, 0, stream>>>(a);
            break;
        case 256:
            if (tiles) march_tiles_kernel<KIND, 256, AGING><<<grid, 256, 0, stream>>>(a);
            else march_batch_kernel<KIND, 256><<<grid, 256, 0, stream>>>(a);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int KIND, bool AGING>
static uint32_t resident_blocks(bool tiles, uint32_t block) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
#define GARG_OCC(B)                                                                                          \
    e = tiles ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, march_tiles_kernel<KIND, B, AGING>, B, 0) \
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, march_batch_kernel<KIND, B>, B, 0)
    if (block == 64) GARG_OCC(64);
    else if (block == 128) GARG_OCC(128);
    else if (block == 256) GARG_OCC(256);
#undef GARG_OCC
    return e == hipSuccess && n > 0 ? (uint32_t)n : 0u;
}

// Workgroups of the batched march one compute unit holds at once: a
// persistent grid beyond that only queues workgroups that find the rays
// already taken.
uint32_t march_batch_resident_blocks(uint32_t kind, bool aging, bool tiles, uint32_t block) {
    switch (kind) {
        case BATCH_SCHWARZSCHILD:
            return aging ? resident_blocks<BATCH_SCHWARZSCHILD, true>(tiles, block)
                         : resident_blocks<BATCH_SCHWARZSCHILD, false>(tiles, block);
        case BATCH_KERR: return resident_blocks<BATCH_KERR, true>(tiles, block);
        case BATCH_KERR_SCHILD: return resident_blocks<BATCH_KERR_SCHILD, true>(tiles, block);
        case BATCH_KERR_POLAR: return resident_blocks<BATCH_KERR_POLAR, true>(tiles, block);
        default: return 0u;
    }
}

hipError_t launch_march_batch(const BatchArgs& a, uint32_t grid, uint32_t block, hipStream_t stream) {
    if (a.refill_threshold == 0u) tile_queue_init_kernel<<<1, 64, 0, stream>>>(a.queue, (uint64_t)grid * block);
    else queue_init_kernel<<<1, 64, 0, stream>>>(a.queue, a.total_rays);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (a.kind) {
        case BATCH_SCHWARZSCHILD:
            return a.aging ? launch_block<BATCH_SCHWARZSCHILD, true>(a, grid, block, stream)
                           : launch_block<BATCH_SCHWARZSCHILD, false>(a, grid, block, stream);
        case BATCH_KERR: return launch_block<BATCH_KERR, true>(a, grid, block, stream);
        case BATCH_KERR_SCHILD: return launch_block<BATCH_KERR_SCHILD, true>(a, grid, block, stream);
        case BATCH_KERR_POLAR: return launch_block<BATCH_KERR_POLAR, true>(a, grid, block, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace garg
