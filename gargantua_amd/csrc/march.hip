// The one-launch-per-frame march (marchImage, march.metal:194-288) and the
// validation probes (probeRays, march.metal:290-341) for gfx950. One thread
// per ray; an 8 x 8 pixel tile is exactly one wave64.
#include "device_march.hpp"
#include "launch.hpp"

namespace garg {

// Shade one camera ray of a pixel with the selected integrator form. The
// result is what marchImage writes to its colour, debug and depth targets.
struct PixelResult {
    f3 color;
    float drift;
    uint32_t steps;
    uint32_t outcome;
    float depth;
};

template <int Form>
GARG_DEV PixelResult shade_pixel(const garg_march_uniforms& u, const Scene& sc, uint32_t px, uint32_t py) {
    PixelResult out;
    f3 dir = camera_direction(u, px, py);
    f3 origin = load3(u.camera_position);
    const uint32_t cap = u.geodesic.step_cap;
    if (u.spin > 0.0f) {
        KerrRayState launched = kerr_launch(u.spin, origin, dir);
        if (Form == GARG_INTEGRATOR_REFERENCE) {
            KerrRefMarch m;
            m.begin(u.spin, launched);
            m.run(u, sc, true, cap);
            out.color = finish_color(m.g, m.falls_in(), m.final_direction(), u, sc);
            out.drift = m.drift;
            out.steps = m.steps;
            out.outcome = m.outcome;
            out.depth = fminf(m.g.depth() / (2.0f * GARG_F(R_ESCAPE)), 1.0f);
        } else if (Form == GARG_INTEGRATOR_KERR_SCHILD) {
            KerrSchildMarch m;
            m.begin(u.spin, launched);
            m.run(u, sc, true, u.compensated != 0u, cap);
            out.color = finish_color(m.g, m.falls_in(), m.final_direction(), u, sc);
            out.drift = m.drift;
            out.steps = m.steps;
            out.outcome = m.outcome;
            out.depth = fminf(m.g.depth() / (2.0f * GARG_F(R_ESCAPE)), 1.0f);
        } else if (Form == GARG_INTEGRATOR_POLAR) {
            KerrHybridMarch m;
            m.begin(u.spin, launched);
            m.run(u, sc, true, u.compensated != 0u, cap);
            out.color = finish_color(m.gather(), m.falls_in(), m.final_direction(), u, sc);
            out.drift = m.max_drift();
            out.steps = m.step_count();
            out.outcome = m.result();
            out.depth = fminf(m.gather().depth() / (2.0f * GARG_F(R_ESCAPE)), 1.0f);
        } else {
            KerrFastMarch m;
            m.begin(u.spin, launched);
            m.run(u, sc, true, u.compensated != 0u, cap);
            out.color = finish_color(m.g, m.falls_in(), m.final_direction(), u, sc);
            out.drift = m.drift;
            out.steps = m.steps;
            out.outcome = m.outcome;
            out.depth = fminf(m.g.depth() / (2.0f * GARG_F(R_ESCAPE)), 1.0f);
        }
        return out;
    }
    Prepared p = prepare_ray(origin, dir, u, sc);
    if (p.sky_only) {
        out.color = u.star_brightness > 0.0f ? starfield(p.sky, u.star_seed, sc.blackbody) * u.star_brightness
                                             : f3{0, 0, 0};
        out.drift = 0.0f;
        out.steps = 0;
        out.outcome = GARG_RAY_ESCAPED;
        out.depth = 1.0f;
        return out;
    }
    SchwMarchT<Form != GARG_INTEGRATOR_REFERENCE> m;
    m.begin(p.ray);
    m.run(u, sc, true, Form != GARG_INTEGRATOR_REFERENCE && u.compensated != 0u, cap);
    bool in = m.falls_in();
    out.color = finish_color(m.g, in, in ? f3{0, 0, 0} : m.sky_direction(u, sc), u, sc);
    out.drift = ray_drift(m.ray);
    out.steps = m.steps;
    out.outcome = m.outcome;
    out.depth = fminf(m.g.depth() / (2.0f * GARG_F(R_ESCAPE)), 1.0f);
    return out;
}

template <int Form>
__global__ __launch_bounds__(64) void march_image_kernel(MarchImageArgs a) {
    const garg_march_uniforms& u = a.uniforms;
    uint32_t gx = blockIdx.x * 8u + (threadIdx.x & 7u);
    uint32_t gy = blockIdx.y * 8u + (threadIdx.x >> 3);
    uint32_t px = u.tile_origin[0] + gx, py = u.tile_origin[1] + gy;
    bool active = gx < a.extent_w && gy < a.extent_h && px < u.resolution[0] && py < u.resolution[1];
    PixelResult r{f3{0, 0, 0}, 0.0f, 0u, GARG_RAY_ESCAPED, 1.0f};
    if (active) {
        r = shade_pixel<Form>(u, a.scene, px, py);
        size_t index = (size_t)py * u.resolution[0] + px;
        store_half4(&a.color[index], r.color.x, r.color.y, r.color.z, 1.0f);
        store_half4(&a.debug[index], r.drift, (float)r.steps, (float)r.outcome, 1.0f);
        if (a.depth) a.depth[index] = r.depth;
    }
    add_counters(a.counters, active, r.drift > u.drift_budget, r.outcome == GARG_RAY_EXHAUSTED,
                 r.outcome == GARG_RAY_CAPTURED, r.steps);
}

// One launch's counters added to a request's counters and a batch's totals:
// a batch that gives a lone request to the one-launch kernel counts it into
// a block of its own, so that the kernel's waves add to one place only.
__global__ void add_counters_kernel(garg_march_counters* a, garg_march_counters* b, const garg_march_counters* src) {
    if (threadIdx.x != 0) return;
    const garg_march_counters c = *src;
    garg_march_counters* const dst[2] = {a, b};
    for (int k = 0; k < 2; ++k) {
        uint32_t* d = reinterpret_cast<uint32_t*>(dst[k]);
        atomicAdd(&d[0], c.rays);
        atomicAdd(&d[1], c.over_drift_budget);
        atomicAdd(&d[2], c.exhausted);
        atomicAdd(&d[3], c.captured);
        atomicAdd((unsigned long long*)&dst[k]->steps, (unsigned long long)c.steps);
    }
}

hipError_t launch_add_counters(garg_march_counters* a, garg_march_counters* b, const garg_march_counters* src,
                               hipStream_t stream) {
    add_counters_kernel<<<1, 64, 0, stream>>>(a, b, src);
    return hipGetLastError();
}

template <int Form>
__global__ __launch_bounds__(64) void probe_kernel(ProbeArgs a) {
    uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= a.count) return;
    const garg_march_uniforms& u = a.uniforms;
    const Scene& sc = a.scene;
    const garg_ray_probe probe = a.probes[i];
    f3 origin = load3(probe.origin), dir = load3(probe.direction);
    garg_ray_probe_result out;
    f3 pos, d;
    if (u.spin > 0.0f) {
        KerrRayState launched = kerr_launch(u.spin, origin, dir);
        if (Form == GARG_INTEGRATOR_REFERENCE) {
            KerrRefMarch m;
            m.begin(u.spin, launched);
            m.run(u, sc, false, u.geodesic.step_cap);
            pos = m.final_position();
            d = m.final_direction();
            out.drift = m.drift;
            out.outcome = m.outcome;
            out.steps = m.steps;
        } else if (Form == GARG_INTEGRATOR_KERR_SCHILD) {
            KerrSchildMarch m;
            m.begin(u.spin, launched);
            m.run(u, sc, false, u.compensated != 0u, u.geodesic.step_cap);
            pos = m.final_position();
            d = m.final_direction();
            out.drift = m.drift;
            out.outcome = m.outcome;
            out.steps = m.steps;
        } else if (Form == GARG_INTEGRATOR_POLAR) {
            KerrHybridMarch m;
            m.begin(u.spin, launched);
            m.run(u, sc, false, u.compensated != 0u, u.geodesic.step_cap);
            pos = m.final_position();
            d = m.final_direction();
            out.drift = m.max_drift();
            out.outcome = m.result();
            out.steps = m.step_count();
        } else {
            KerrFastMarch m;
            m.begin(u.spin, launched);
            m.run(u, sc, false, u.compensated != 0u, u.geodesic.step_cap);
            pos = m.final_position();
            d = m.final_direction();
            out.drift = m.drift;
            out.outcome = m.outcome;
            out.steps = m.steps;
        }
        out.energy = launched.energy;
        out.angular_momentum = launched.angular_momentum;
    } else {
        Prepared p = prepare_ray(origin, dir, u, sc);
        if (p.sky_only) {
            pos = ray_position(p.ray);
            d = p.sky;
            out.drift = 0.0f;
            out.outcome = GARG_RAY_ESCAPED;
            out.steps = 0;
            out.energy = p.ray.energy;
            out.angular_momentum = p.ray.angular_momentum;
        } else {
            SchwMarchT<Form != GARG_INTEGRATOR_REFERENCE> m;
            m.begin(p.ray);
            m.run(u, sc, false, Form != GARG_INTEGRATOR_REFERENCE && u.compensated != 0u, u.geodesic.step_cap);
            pos = m.position();
            d = (m.outcome == GARG_RAY_ESCAPED && !m.falls_in()) ? m.sky_direction(u, sc) : m.direction();
            out.drift = ray_drift(m.ray);
            out.outcome = m.outcome;
            out.steps = m.steps;
            out.energy = ray_energy(m.ray.s);
            out.angular_momentum = ray_angular_momentum(m.ray.s);
        }
    }
    out.position[0] = pos.x;
    out.position[1] = pos.y;
    out.position[2] = pos.z;
    out.direction[0] = d.x;
    out.direction[1] = d.y;
    out.direction[2] = d.z;
    a.results[i] = out;
}

hipError_t launch_march_image(const MarchImageArgs& a, uint32_t integrator, hipStream_t stream) {
    dim3 grid((a.extent_w + 7u) / 8u, (a.extent_h + 7u) / 8u);
    if (integrator == GARG_INTEGRATOR_REFERENCE)
        march_image_kernel<GARG_INTEGRATOR_REFERENCE><<<grid, 64, 0, stream>>>(a);
    else if (integrator == GARG_INTEGRATOR_KERR_SCHILD)
        march_image_kernel<GARG_INTEGRATOR_KERR_SCHILD><<<grid, 64, 0, stream>>>(a);
    else if (integrator == GARG_INTEGRATOR_POLAR)
        march_image_kernel<GARG_INTEGRATOR_POLAR><<<grid, 64, 0, stream>>>(a);
    else
        march_image_kernel<GARG_INTEGRATOR_FAST><<<grid, 64, 0, stream>>>(a);
    return hipGetLastError();
}

template <int Form>
static uint32_t image_resident_blocks() {
    int n = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, march_image_kernel<Form>, 64, 0) == hipSuccess && n > 0
               ? (uint32_t)n
               : 0u;
}

uint32_t march_image_resident_blocks(uint32_t integrator) {
    switch (integrator) {
        case GARG_INTEGRATOR_REFERENCE: return image_resident_blocks<GARG_INTEGRATOR_REFERENCE>();
        case GARG_INTEGRATOR_KERR_SCHILD: return image_resident_blocks<GARG_INTEGRATOR_KERR_SCHILD>();
        case GARG_INTEGRATOR_POLAR: return image_resident_blocks<GARG_INTEGRATOR_POLAR>();
        default: return image_resident_blocks<GARG_INTEGRATOR_FAST>();
    }
}

hipError_t launch_probe(const ProbeArgs& a, uint32_t integrator, hipStream_t stream) {
    dim3 grid((a.count + 63u) / 64u);
    if (integrator == GARG_INTEGRATOR_REFERENCE)
        probe_kernel<GARG_INTEGRATOR_REFERENCE><<<grid, 64, 0, stream>>>(a);
    else if (integrator == GARG_INTEGRATOR_KERR_SCHILD)
        probe_kernel<GARG_INTEGRATOR_KERR_SCHILD><<<grid, 64, 0, stream>>>(a);
    else if (integrator == GARG_INTEGRATOR_POLAR)
        probe_kernel<GARG_INTEGRATOR_POLAR><<<grid, 64, 0, stream>>>(a);
    else
        probe_kernel<GARG_INTEGRATOR_FAST><<<grid, 64, 0, stream>>>(a);
    return hipGetLastError();
}

// The FP32 ceilings the march is measured against: every thread runs
// `iterations` rounds of 32 independent non-packed fused multiply-adds
// (packed FP32 instructions are disabled in this build) and stores their
// sum, so that none of them is dead code. The operand form sets the issue
// rate on gfx950: FORM 0 reads two VGPRs and a literal (v_fmamk_f32), FORM 1
// three VGPRs (v_fma_f32), FORM 2 a VGPR and two SGPRs (the wave-uniform
// kernel arguments).
constexpr int VALU_RATE_CHAINS = 32;

template <int FORM>
__global__ __launch_bounds__(256) void valu_rate_kernel(float* out, uint32_t threads, uint32_t iterations, float m,
                                                       float c) {
    uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= threads) return;
    // Per-lane multiplier and addend (VGPRs) for FORM 1, close to m and c.
    const float mv = m + (float)(t & 7u) * 1e-9f, cv = c + (float)(t & 3u) * 1e-12f;
    float x[VALU_RATE_CHAINS];
#pragma unroll
    for (int j = 0; j < VALU_RATE_CHAINS; ++j) x[j] = (float)(t & 1023u) * 1e-3f + (float)j;
#pragma unroll 2
    for (uint32_t i = 0; i < iterations; ++i) {
#pragma unroll
        for (int j = 0; j < VALU_RATE_CHAINS; ++j) {
            if constexpr (FORM == 0) x[j] = __builtin_fmaf(x[j], 0.99999994f, cv);
            else if constexpr (FORM == 1) x[j] = __builtin_fmaf(x[j], mv, cv);
            else x[j] = __builtin_fmaf(x[j], m, c);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < VALU_RATE_CHAINS; ++j) s += x[j];
    out[t] = s;
}

hipError_t launch_valu_rate(float* out, uint32_t form, uint32_t threads, uint32_t iterations, hipStream_t stream) {
    const dim3 grid((threads + 255u) / 256u);
    if (form == 0) valu_rate_kernel<0><<<grid, 256, 0, stream>>>(out, threads, iterations, 0.99999994f, 1e-7f);
    else if (form == 1) valu_rate_kernel<1><<<grid, 256, 0, stream>>>(out, threads, iterations, 0.99999994f, 1e-7f);
    else valu_rate_kernel<2><<<grid, 256, 0, stream>>>(out, threads, iterations, 0.99999994f, 1e-7f);
    return hipGetLastError();
}

}  // namespace garg
