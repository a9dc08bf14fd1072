// Learned smoothing-length proposal: a 4 -> 32 -> 32 -> 1 SiLU MLP with FP32
// weights and BF16 inputs, trained with Adam by the kernels below.
//
// Parameter layout (1249 floats): W1[32][4], b1[32], W2[32][32], b2[32],
// W3[32], b3.  The network maps the normalised features of a gas particle
// (k_mlp_features) to a correction of log d_k: log d_k ~ log d_k(prev) + y.
//
// Training is deterministic: each block of 128 samples forms its gradient
// (per-sample backward in registers, weight gradients as sums over the
// block's samples from LDS) and writes it as one partial; one kernel sums
// the partials in block order and applies Adam.
#include "gf_mlp.h"

namespace {

constexpr int H = 32;
constexpr int NIN = 4;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + H * NIN;
constexpr int OFF_W2 = OFF_B1 + H;
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + H;
static_assert(OFF_B3 + 1 == GF_MLP_PARAMS, "parameter layout");

constexpr int FB = 256;      // forward block
constexpr int GB = 128;      // gradient block (samples per partial)

__device__ __forceinline__ float bf16_to_f32(uint16_t v) {
    return __uint_as_float(uint32_t(v) << 16);
}

// round to nearest even
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);   // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + expf(-z)); }

// Forward pass of one sample; a1, a2 receive the activations, z1, z2 the
// pre-activations (for the backward pass).
__device__ __forceinline__ float mlp_forward(const float* __restrict__ p, const float x[NIN],
                                             float z1[H], float a1[H], float z2[H], float a2[H]) {
#pragma unroll
    for (int i = 0; i < H; ++i) {
        float s = p[OFF_B1 + i];
#pragma unroll
        for (int k = 0; k < NIN; ++k) s = fmaf(p[OFF_W1 + i * NIN + k], x[k], s);
        z1[i] = s;
        a1[i] = s * sigmoidf_(s);
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
        float s = p[OFF_B2 + i];
#pragma unroll
        for (int k = 0; k < H; ++k) s = fmaf(p[OFF_W2 + i * H + k], a1[k], s);
        z2[i] = s;
        a2[i] = s * sigmoidf_(s);
    }
    float y = p[OFF_B3];
#pragma unroll
    for (int k = 0; k < H; ++k) y = fmaf(p[OFF_W3 + k], a2[k], y);
    return y;
}

__global__ __launch_bounds__(FB) void k_mlp_forward(int n, const uint16_t* __restrict__ x,
                                                    const float* __restrict__ params,
                                                    float* __restrict__ out) {
    __shared__ float p[GF_MLP_PARAMS];
    for (int i = threadIdx.x; i < GF_MLP_PARAMS; i += FB) p[i] = params[i];
    __syncthreads();
    const int t = blockIdx.x * FB + threadIdx.x;
    if (t >= n) return;
    float xi[NIN], z1[H], a1[H], z2[H], a2[H];
#pragma unroll
    for (int k = 0; k < NIN; ++k) xi[k] = bf16_to_f32(x[NIN * size_t(t) + k]);
    out[t] = mlp_forward(p, xi, z1, a1, z2, a2);
}

// Gradient over the block's samples -> partial[block][GF_MLP_PARAMS] of
//   tau <= 0:  the mean squared error 0.5 (y - target)^2 / n_total
//   tau in (0, 1): the mean pinball loss of the tau-quantile,
//              max(tau (t - y), (1 - tau) (y - t)) / n_total;
// loss_part[block] gets the block's sum of squared errors, or of pinball
// losses.
__global__ __launch_bounds__(GB) void k_mlp_grad(int n, int n_total, const uint16_t* __restrict__ x,
                                                 const float* __restrict__ target,
                                                 const float* __restrict__ params, float tau,
                                                 float* __restrict__ partial,
                                                 float* __restrict__ loss_part) {
    __shared__ float p[GF_MLP_PARAMS];
    __shared__ float s_a1[GB][H + 1];
    __shared__ float s_dz2[GB][H + 1];
    __shared__ float s_dz1[GB][H + 1];
    __shared__ float s_a2[GB][H + 1];
    __shared__ float s_x[GB][NIN];
    __shared__ float s_dy[GB];
    __shared__ float s_loss[GB];
    for (int i = threadIdx.x; i < GF_MLP_PARAMS; i += GB) p[i] = params[i];
    __syncthreads();
    const int tid = threadIdx.x;
    const int t = blockIdx.x * GB + tid;
    const bool ok = t < n;
    float xi[NIN], z1[H], a1[H], z2[H], a2[H];
#pragma unroll
    for (int k = 0; k < NIN; ++k) xi[k] = ok ? bf16_to_f32(x[NIN * size_t(t) + k]) : 0.0f;
    const float y = mlp_forward(p, xi, z1, a1, z2, a2);
    const float err = ok ? y - target[t] : 0.0f;
    float dy, lval;
    if (tau > 0.0f) {
        dy = ok ? (err >= 0.0f ? 1.0f - tau : -tau) / float(n_total) : 0.0f;
        lval = err >= 0.0f ? (1.0f - tau) * err : -tau * err;
    } else {
        dy = err / float(n_total);
        lval = err * err;
    }
    // backward through layer 3 and the SiLUs
    float dz2[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const float sg = sigmoidf_(z2[i]);
        dz2[i] = dy * p[OFF_W3 + i] * sg * (1.0f + z2[i] * (1.0f - sg));
    }
#pragma unroll
    for (int k = 0; k < H; ++k) {
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < H; ++i) s = fmaf(p[OFF_W2 + i * H + k], dz2[i], s);
        const float sg = sigmoidf_(z1[k]);
        s_dz1[tid][k] = s * sg * (1.0f + z1[k] * (1.0f - sg));
        s_a1[tid][k] = a1[k];
        s_dz2[tid][k] = dz2[k];
        s_a2[tid][k] = a2[k];
    }
#pragma unroll
    for (int k = 0; k < NIN; ++k) s_x[tid][k] = xi[k];
    s_dy[tid] = dy;
    s_loss[tid] = lval;
    __syncthreads();
    // weight gradients: sums over the block's samples, one output per loop step
    float* out = partial + size_t(blockIdx.x) * GF_MLP_PARAMS;
    for (int q = tid; q < GF_MLP_PARAMS; q += GB) {
        float s = 0.0f;
        if (q < OFF_B1) {                       // W1[i][k]
            const int i = q / NIN, k = q % NIN;
            for (int j = 0; j < GB; ++j) s = fmaf(s_dz1[j][i], s_x[j][k], s);
        } else if (q < OFF_W2) {                // b1[i]
            const int i = q - OFF_B1;
            for (int j = 0; j < GB; ++j) s += s_dz1[j][i];
        } else if (q < OFF_B2) {                // W2[i][k]
            const int i = (q - OFF_W2) / H, k = (q - OFF_W2) % H;
            for (int j = 0; j < GB; ++j) s = fmaf(s_dz2[j][i], s_a1[j][k], s);
        } else if (q < OFF_W3) {                // b2[i]
            const int i = q - OFF_B2;
            for (int j = 0; j < GB; ++j) s += s_dz2[j][i];
        } else if (q < OFF_B3) {                // W3[k]
            const int k = q - OFF_W3;
            for (int j = 0; j < GB; ++j) s = fmaf(s_dy[j], s_a2[j][k], s);
        } else {                                // b3
            for (int j = 0; j < GB; ++j) s += s_dy[j];
        }
        out[q] = s;
    }
    if (tid == 0) {
        float l = 0.0f;
        for (int j = 0; j < GB; ++j) l += s_loss[j];
        loss_part[blockIdx.x] = l;
    }
}

// Sums the partials in block order; optionally applies one Adam step.
__global__ __launch_bounds__(256) void k_mlp_reduce_adam(int nblocks, const float* __restrict__ partial,
                                                         const float* __restrict__ loss_part,
                                                         float* __restrict__ grad,
                                                         float* __restrict__ loss, float* params,
                                                         float* m, float* v, int step, float lr,
                                                         float b1, float b2, float eps) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q == 0 && loss) {
        float l = 0.0f;
        for (int b = 0; b < nblocks; ++b) l += loss_part[b];
        *loss = l;
    }
    if (q >= GF_MLP_PARAMS) return;
    float g = 0.0f;
    for (int b = 0; b < nblocks; ++b) g += partial[size_t(b) * GF_MLP_PARAMS + q];
    if (grad) grad[q] = g;
    if (step > 0) {
        const float mq = b1 * m[q] + (1.0f - b1) * g;
        const float vq = b2 * v[q] + (1.0f - b2) * g * g;
        m[q] = mq;
        v[q] = vq;
        const float mh = mq / (1.0f - powf(b1, float(step)));
        const float vh = vq / (1.0f - powf(b2, float(step)));
        params[q] -= lr * mh / (sqrtf(vh) + eps);
    }
}

// Features of gas particle i (input order), all free of the length and time
// scales so that one network serves every resolution:
//   x0 = log(d_k(prev) / r_cell), r_cell = (3 k / (4 pi n_cell))^(1/3) the
//        radius that holds k particles at the density of the smallest gas-tree
//        cell with >= k particles (the plain estimate);
//   x1 = log(cell count / k)        (how coarse that estimate is);
//   x2 = div v(prev) * dt           (compression over the step);
//   x3 = log(dt / dt_Courant(prev));
// normalised with (mean, std) and rounded to BF16.
__global__ __launch_bounds__(FB) void k_mlp_features(int n, int k, const double* __restrict__ dk_prev,
                                                     const double* __restrict__ cell_rho,
                                                     const double* __restrict__ cell_count,
                                                     const double* __restrict__ divv,
                                                     const double* __restrict__ dtc_prev, float dt,
                                                     const float* __restrict__ norm,
                                                     uint16_t* __restrict__ x,
                                                     float* __restrict__ raw) {
    const int i = blockIdx.x * FB + threadIdx.x;
    if (i >= n) return;
    const double r_cell = cbrt(3.0 * k / (4.0 * M_PI * fmax(cell_rho[i], 1e-300)));
    float f[NIN];
    f[0] = float(log(dk_prev[i] / r_cell));
    f[1] = float(log(fmax(cell_count[i], 1.0) / double(k)));
    f[2] = float(divv[i] * double(dt));
    f[3] = float(log(fmax(double(dt), 1e-300) / fmax(dtc_prev[i], 1e-300)));
#pragma unroll
    for (int q = 0; q < NIN; ++q) {
        if (raw) raw[NIN * size_t(i) + q] = f[q];
        x[NIN * size_t(i) + q] = f32_to_bf16((f[q] - norm[q]) / norm[NIN + q]);
    }
}

// First search radius: r = d_k(prev) exp(y + margin).
__global__ __launch_bounds__(FB) void k_mlp_propose(int n, const uint16_t* __restrict__ x,
                                                    const float* __restrict__ params,
                                                    const double* __restrict__ dk_prev,
                                                    float margin, double* __restrict__ r) {
    __shared__ float p[GF_MLP_PARAMS];
    for (int i = threadIdx.x; i < GF_MLP_PARAMS; i += FB) p[i] = params[i];
    __syncthreads();
    const int t = blockIdx.x * FB + threadIdx.x;
    if (t >= n) return;
    float xi[NIN], z1[H], a1[H], z2[H], a2[H];
#pragma unroll
    for (int k = 0; k < NIN; ++k) xi[k] = bf16_to_f32(x[NIN * size_t(t) + k]);
    const float y = mlp_forward(p, xi, z1, a1, z2, a2);
    r[t] = dk_prev[t] * exp(double(y) + double(margin));
}

}  // namespace

size_t gf_mlp_partial_floats(int n) {
    return size_t(gf_div_up(std::max(n, 1), GB)) * (GF_MLP_PARAMS + 1);
}

int gf_mlp_forward(int n, const uint16_t* x, const float* params, float* out, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_mlp_forward, dim3(gf_div_up(n, FB)), dim3(FB), 0, s, n, x, params, out);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_mlp_grad(int n, int n_total, const uint16_t* x, const float* target, const float* params,
                float tau, float* grad, float* loss, float* partial, hipStream_t s) {
    if (n <= 0) return 0;
    const int nb = gf_div_up(n, GB);
    float* loss_part = partial + size_t(nb) * GF_MLP_PARAMS;
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(GB), 0, s, n, n_total, x, target, params, tau,
                       partial, loss_part);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(gf_div_up(GF_MLP_PARAMS, 256)), dim3(256), 0, s, nb,
                       partial, loss_part, grad, loss, nullptr, nullptr, nullptr, 0, 0.f, 0.f, 0.f,
                       0.f);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_mlp_train_step(int n, const uint16_t* x, const float* target, float* params, float* m,
                      float* v, int step, float lr, float b1, float b2, float eps, float tau,
                      float* loss, float* partial, hipStream_t s) {
    if (n <= 0) return 0;
    const int nb = gf_div_up(n, GB);
    float* loss_part = partial + size_t(nb) * GF_MLP_PARAMS;
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(GB), 0, s, n, n, x, target, params, tau,
                       partial, loss_part);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(gf_div_up(GF_MLP_PARAMS, 256)), dim3(256), 0, s, nb,
                       partial, loss_part, nullptr, loss, params, m, v, step, lr, b1, b2, eps);
    GF_LAUNCH_CHECK();
    return 0;
}

// ------------------------------------------- training step without allocation
// The partial gradients and losses of gf_mlp_train_step_noalloc live in this
// file-scope array, sized for GF_MLP_NOALLOC_MAX samples (512 partials of
// GF_MLP_PARAMS + 1 floats, 2.56 MB per device, held until the process
// ends).  In host code `scratch` names a host-side stand-in, not device
// memory: kernels receive the device address that hipGetSymbolAddress
// returns, looked up once per device on that device's first call and kept.
// Each call overwrites the partials of its own blocks before it reads them,
// so nothing has to start at zero; calls that overlap on one device race.
static __device__ __align__(256) float scratch[(GF_MLP_NOALLOC_MAX / GB) * (GF_MLP_PARAMS + 1)];
static_assert(GF_MLP_NOALLOC_MAX % GB == 0, "whole gradient blocks");

namespace {
constexpr int kMaxDevices = 64;
float* g_scratch_dev[kMaxDevices] = {};
}  // namespace

int gf_mlp_train_step_noalloc(int n, const uint16_t* x, const float* target, float* params,
                              float* m, float* v, int step, float lr, float b1, float b2,
                              float eps, float tau, float* loss, hipStream_t s) {
    if (n > GF_MLP_NOALLOC_MAX) {
        gf_set_error("gf_mlp_train_step_noalloc_dev: %d samples exceed its scratch of %d", n,
                     GF_MLP_NOALLOC_MAX);
        return -1;
    }
    if (n <= 0) return 0;
    int dev = 0;
    GF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        gf_set_error("gf_mlp_train_step_noalloc_dev: device %d beyond the %d supported", dev,
                     kMaxDevices);
        return -1;
    }
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        GF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            gf_set_error("gf_mlp_train_step_noalloc_dev: no device address for its scratch");
            return -1;
        }
        g_scratch_dev[dev] = static_cast<float*>(p);
    }
    static_assert(sizeof(scratch) / sizeof(float) >=
                      size_t(GF_MLP_NOALLOC_MAX / GB) * (GF_MLP_PARAMS + 1),
                  "scratch holds the partials of GF_MLP_NOALLOC_MAX samples");
    return gf_mlp_train_step(n, x, target, params, m, v, step, lr, b1, b2, eps, tau, loss,
                             g_scratch_dev[dev], s);
}

int gf_mlp_features(int n, int k, const double* dk_prev, const double* cell_rho,
                    const double* cell_count, const double* divv, const double* dtc_prev,
                    float dt, const float* norm, uint16_t* x, float* raw, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_mlp_features, dim3(gf_div_up(n, FB)), dim3(FB), 0, s, n, k, dk_prev,
                       cell_rho, cell_count, divv, dtc_prev, dt, norm, x, raw);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_mlp_propose(int n, const uint16_t* x, const float* params, const double* dk_prev,
                   float margin, double* r, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_mlp_propose, dim3(gf_div_up(n, FB)), dim3(FB), 0, s, n, x, params,
                       dk_prev, margin, r);
    GF_LAUNCH_CHECK();
    return 0;
}
