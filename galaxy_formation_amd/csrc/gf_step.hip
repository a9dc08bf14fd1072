// The caller's side of the force path on the device: the force sum and
// time-step rule of galaxy_formation/simulation.py:61-72, and the drift, rung
// and kick steps of a block leapfrog (simulation.py:114-149 widened to
// several rungs).  One thread per particle; FP64 without contraction, so
// each result is the expression as those lines write it.
#include "gf_step.h"

namespace {

constexpr int TB = 256;

__global__ void k_reset_bits(unsigned long long* bits) {
    if (threadIdx.x == 0) *bits = 0ull;
}

// simulation.py:68-71.  |acc| of non-negative doubles orders like their bit
// patterns, so the largest is an atomic maximum of the bits, one per block.
__global__ __launch_bounds__(TB) void k_combine(int n, int ng, double eps,
                                                const double* __restrict__ ag,
                                                const double* __restrict__ as,
                                                const double* __restrict__ dtc,
                                                double* __restrict__ acc,
                                                double* __restrict__ dt_ind,
                                                unsigned long long* __restrict__ amax_bits) {
#pragma clang fp contract(off)
    __shared__ double s_max[TB / 64];
    const int i = blockIdx.x * TB + threadIdx.x;
    double amag = 0.0;
    if (i < n) {
        double ax = ag[3 * size_t(i)], ay = ag[3 * size_t(i) + 1], az = ag[3 * size_t(i) + 2];
        if (i < ng) {
            ax = ax + as[3 * size_t(i)];
            ay = ay + as[3 * size_t(i) + 1];
            az = az + as[3 * size_t(i) + 2];
        }
        acc[3 * size_t(i)] = ax;
        acc[3 * size_t(i) + 1] = ay;
        acc[3 * size_t(i) + 2] = az;
        amag = sqrt(ax * ax + ay * ay + az * az);
        double d = 0.3 * sqrt(eps / fmax(amag, 1e-30));
        if (i < ng) d = fmin(d, dtc[i]);
        dt_ind[i] = d;
    }
    const double wm = gf_wave_max_d(amag);
    if (__lane_id() == 0) s_max[threadIdx.x >> 6] = wm;
    __syncthreads();
    if (threadIdx.x == 0) {
        double bm = s_max[0];
        for (int w = 1; w < TB / 64; ++w) bm = fmax(bm, s_max[w]);
        atomicMax(amax_bits, (unsigned long long)__double_as_longlong(bm));
    }
}

// dt_grav = 0.3 sqrt(eps / max(max |acc|, 1e-30)), in place of the bits
__global__ void k_dt_grav(double eps, unsigned long long* bits) {
#pragma clang fp contract(off)
    if (threadIdx.x == 0) {
        const double amax = __longlong_as_double((long long)*bits);
        reinterpret_cast<double*>(bits)[0] = 0.3 * sqrt(eps / fmax(amax, 1e-30));
    }
}

__global__ __launch_bounds__(TB) void k_step_drift(int n, double dt, long long k, int rungs,
                                                   double* __restrict__ pos,
                                                   const double* __restrict__ vel,
                                                   const int64_t* __restrict__ rung,
                                                   uint8_t* __restrict__ active,
                                                   int* __restrict__ count) {
#pragma clang fp contract(off)
    __shared__ int s_cnt[TB / 64];
    const int i = blockIdx.x * TB + threadIdx.x;
    bool a = false;
    if (i < n) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const size_t o = 3 * size_t(i) + c;
            pos[o] = pos[o] + dt * vel[o];
        }
        const long long period = 1ll << (rungs - 1 - int(rung[i]));
        a = k % period == 0;
        active[i] = a ? 1 : 0;
    }
    const int wc = __popcll(__ballot(a));
    if (__lane_id() == 0) s_cnt[threadIdx.x >> 6] = wc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int bc = 0;
        for (int w = 0; w < TB / 64; ++w) bc += s_cnt[w];
        if (bc) atomicAdd(count, bc);
    }
}

__device__ __forceinline__ int rung_of(double dt_max, double dt, int rungs) {
    const double r = ceil(log2(fmax(dt_max / dt, 1.0)));
    return int(fmin(fmax(r, 0.0), double(rungs - 1)));
}

__global__ __launch_bounds__(TB) void k_step_kick(int n, int ng, long long k, int rungs,
                                                  double dt_max, int mode,
                                                  const uint8_t* __restrict__ active,
                                                  const double* __restrict__ acc,
                                                  const double* __restrict__ dudt,
                                                  const double* __restrict__ dt_ind,
                                                  int64_t* __restrict__ rung,
                                                  double* __restrict__ vel, double* __restrict__ u,
                                                  double u_floor) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const bool act = active[i] != 0;
    int r = int(rung[i]);
    if (mode == 2 || (mode == 1 && act)) {
        int nr = rung_of(dt_max, dt_ind[i], rungs);
        // a longer step only where this sub-step is also its boundary
        if (mode == 1 && nr < r && k % (1ll << (rungs - 1 - nr)) != 0) nr = r;
        r = nr;
        rung[i] = r;
    }
    if (!act) return;
    const double h = 0.5 * (dt_max / double(1ll << r));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t o = 3 * size_t(i) + c;
        vel[o] = vel[o] + h * acc[o];
    }
    if (i < ng) u[i] = fmax(u[i] + h * dudt[i], u_floor);
}

}  // namespace

int gf_combine(int n, int ng, double eps, const double* acc_grav, const double* acc_sph,
               const double* dt_c, double* acc, double* dt_ind, double* dt_grav, hipStream_t s) {
    unsigned long long* bits = reinterpret_cast<unsigned long long*>(dt_grav);
    hipLaunchKernelGGL(k_reset_bits, dim3(1), dim3(64), 0, s, bits);
    if (n > 0)
        hipLaunchKernelGGL(k_combine, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, ng, eps, acc_grav,
                           acc_sph, dt_c, acc, dt_ind, bits);
    hipLaunchKernelGGL(k_dt_grav, dim3(1), dim3(64), 0, s, eps, bits);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_step_drift(int n, double dt, long long k, int rungs, double* pos, const double* vel,
                  const int64_t* rung, uint8_t* active, int* count, hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_step_drift, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, dt, k, rungs, pos,
                       vel, rung, active, count);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_step_kick(int n, int ng, long long k, int rungs, double dt_max, int mode,
                 const uint8_t* active, const double* acc, const double* dudt,
                 const double* dt_ind, int64_t* rung, double* vel, double* u, double u_floor,
                 hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_step_kick, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, ng, k, rungs, dt_max,
                       mode, active, acc, dudt, dt_ind, rung, vel, u, u_floor);
    GF_LAUNCH_CHECK();
    return 0;
}
