// Rendering entry points of the reference library on the device: the
// orthographic particle deposit of nbody_render (cpp/nbody.cpp:445-469) and
// the per-pixel tone map of nbody_tonemap (cpp/nbody.cpp:474-503).
//
// Deposit: one thread per particle projects it and adds its four bilinear
// weights with FP32 atomics.  The projection parameters are read once per
// block into LDS.  Sums in a pixel are order-dependent, so they match a
// sequential deposit to FP32 rounding.
//
// Tone map: a 2D grid of 64 x 4 threads, one pixel each, a row of 64
// consecutive pixels per wave, so image reads of every group plane are
// coalesced.  The glow map's sample row is shared by the whole row of
// threads; the float expressions are the reference's, so the bytes agree to
// the last bit where the compilers round alike and within one unit
// otherwise.
#include "gf_render.h"

namespace {

constexpr int kSplatBlock = 256;
constexpr int kShadeX = 64, kShadeY = 4;

// the centre and the two screen rows of the rotation
struct View {
    double c[3];
    double ax[3];
    double ay[3];
};

__global__ __launch_bounds__(kSplatBlock) void k_splat(int n, const double* __restrict__ pos,
                                                       const uint8_t* __restrict__ group,
                                                       const double* __restrict__ rot,
                                                       const double* __restrict__ center,
                                                       double scale, int W, int H,
                                                       float* __restrict__ img) {
    __shared__ View view;
    if (threadIdx.x < 3) {
        view.c[threadIdx.x] = center[threadIdx.x];
        view.ax[threadIdx.x] = rot[threadIdx.x];
        view.ay[threadIdx.x] = rot[3 + threadIdx.x];
    }
    __syncthreads();
    const int i = blockIdx.x * kSplatBlock + threadIdx.x;
    if (i >= n) return;
    const double dx = pos[3 * size_t(i)] - view.c[0];
    const double dy = pos[3 * size_t(i) + 1] - view.c[1];
    const double dz = pos[3 * size_t(i) + 2] - view.c[2];
    // screen coordinates with pixel centres on integers, y pointing down
    const double sx = 0.5 * W + scale * (view.ax[0] * dx + view.ax[1] * dy + view.ax[2] * dz) - 0.5;
    const double sy = 0.5 * H - scale * (view.ay[0] * dx + view.ay[1] * dy + view.ay[2] * dz) - 0.5;
    const bool inside = sx >= 0 && sy >= 0 && sx < W - 1 && sy < H - 1;   // NaN fails too
    if (!inside) return;
    const int px = int(sx), py = int(sy);
    const float wx = float(sx - px), wy = float(sy - py);
    const float ux = 1 - wx, uy = 1 - wy;
    float* cell = img + size_t(group[i]) * size_t(W) * H + size_t(py) * W + px;
    atomicAdd(cell, ux * uy);
    atomicAdd(cell + 1, wx * uy);
    atomicAdd(cell + W, ux * wy);
    atomicAdd(cell + W + 1, wx * wy);
}

// Sample position of output pixel i in a glow map of m cells along an axis
// of n pixels: the pixel centre mapped to cell centres, clamped inside the
// map so that the bilinear stencil's far cell exists.
__device__ __forceinline__ void glow_axis(int i, float cells_per_pixel, int m, int* cell,
                                          float* frac) {
    const float g = fminf(fmaxf((i + 0.5f) * cells_per_pixel - 0.5f, 0.0f), m - 1.001f);
    *cell = int(g);
    *frac = g - *cell;
}

__device__ __forceinline__ uint8_t saturate_to_byte(float s, float exposure) {
    return uint8_t(255.0f * (1 - expf(-exposure * s)) + 0.5f);
}

__global__ __launch_bounds__(kShadeX* kShadeY) void k_shade(int W, int H, int ngroups,
                                                          const float* __restrict__ img,
                                                          const float* __restrict__ glow, int gw,
                                                          int gh, const float* __restrict__ colors,
                                                          float inv_ref, float K, float exposure,
                                                          uint8_t* __restrict__ out) {
    const int x = blockIdx.x * kShadeX + threadIdx.x;
    const int y = blockIdx.y * kShadeY + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t plane = size_t(W) * H, gplane = size_t(gw) * gh;
    const float norm = 1.0f / log1pf(K);
    int cx, cy;
    float fx, fy;
    glow_axis(x, float(gw) / W, gw, &cx, &fx);
    glow_axis(y, float(gh) / H, gh, &cy, &fy);
    float rgb[3] = {0.0f, 0.0f, 0.0f};
    const size_t pix = size_t(y) * W + x;
    for (int g = 0; g < ngroups; ++g) {
        float level = log1pf(K * inv_ref * img[g * plane + pix]) * norm;
        const float* s = glow + g * gplane + size_t(cy) * gw + cx;
        level += (1 - fy) * ((1 - fx) * s[0] + fx * s[1]) + fy * ((1 - fx) * s[gw] + fx * s[gw + 1]);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) rgb[ch] += level * colors[3 * g + ch];
    }
    uint8_t* o = out + 3 * pix;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) o[ch] = saturate_to_byte(rgb[ch], exposure);
}

}  // namespace

int gf_render_deposit(int n, const double* pos, const uint8_t* group, int ngroups,
                      const double* rot, const double* center, double scale, int W, int H,
                      float* img, hipStream_t s) {
    GF_HIP(hipMemsetAsync(img, 0, sizeof(float) * size_t(ngroups) * W * H, s));
    if (n > 0)
        hipLaunchKernelGGL(k_splat, dim3(gf_div_up(n, kSplatBlock)), dim3(kSplatBlock), 0, s, n,
                           pos, group, rot, center, scale, W, H, img);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_render_tonemap(int W, int H, int ngroups, const float* img, const float* glow, int gw,
                      int gh, const float* colors, float inv_ref, float K, float exposure,
                      uint8_t* out, hipStream_t s) {
    if (W > 0 && H > 0)
        hipLaunchKernelGGL(k_shade, dim3(gf_div_up(W, kShadeX), gf_div_up(H, kShadeY)),
                           dim3(kShadeX, kShadeY), 0, s, W, H, ngroups, img, glow, gw, gh, colors,
                           inv_ref, K, exposure, out);
    GF_LAUNCH_CHECK();
    return 0;
}
