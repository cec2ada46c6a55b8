// Present for gfx950: exposure, the Khronos PBR Neutral tone curve
// (present.metal:19-35) and sRGB encoding into 8-bit RGBA, or the drift heat
// map debug view (present.metal:40-46). One thread per pixel.
#include "device_common.hpp"
#include "launch.hpp"

namespace garg {

GARG_DEV f3 pbr_neutral(f3 c) {
    const float start = 0.8f - 0.04f;
    const float desaturation = 0.15f;
    float x = fminf(c.x, fminf(c.y, c.z));
    float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
    c = c - f3{offset, offset, offset};
    float peak = fmaxf(c.x, fmaxf(c.y, c.z));
    if (peak < start) return c;
    const float d = 1.0f - start;
    float new_peak = 1.0f - d * d / (peak + d - start);
    c = c * (new_peak / peak);
    float g = 1.0f - 1.0f / (desaturation * (peak - new_peak) + 1.0f);
    return c + (f3{new_peak, new_peak, new_peak} - c) * g;
}

GARG_DEV float srgb_encode(float x) {
    x = fminf(fmaxf(x, 0.0f), 1.0f);
    return x <= 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}

GARG_DEV uint8_t to_byte(float x) { return (uint8_t)(fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f + 0.5f); }

__global__ __launch_bounds__(256) void tonemap_kernel(TonemapArgs a) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.pixels) return;
    f3 out;
    if (a.debug_view) {
        h4 d = a.debug[i];
        float t = fminf(fmaxf((float)d.x / a.drift_budget, 0.0f), 2.0f) * 0.5f;
        f3 heat{t, 1.0f - fabsf(t - 0.5f) * 2.0f, 1.0f - t};
        float shade = 0.2f + 0.8f * fminf(fmaxf((float)d.y / (float)STEP_CAP_INTERACTIVE, 0.0f), 1.0f);
        if ((float)d.z < 0.5f) shade *= 0.35f;
        out = heat * shade;
    } else {
        h4 c = a.color[i];
        f3 color{fmaxf((float)c.x * a.exposure, 0.0f), fmaxf((float)c.y * a.exposure, 0.0f),
                 fmaxf((float)c.z * a.exposure, 0.0f)};
        out = pbr_neutral(color);
    }
    uchar4 px;
    px.x = to_byte(srgb_encode(out.x));
    px.y = to_byte(srgb_encode(out.y));
    px.z = to_byte(srgb_encode(out.z));
    px.w = 255u;
    reinterpret_cast<uchar4*>(a.out)[i] = px;
}

hipError_t launch_tonemap(const TonemapArgs& a, hipStream_t stream) {
    tonemap_kernel<<<dim3((a.pixels + 255u) / 256u), 256, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace garg
