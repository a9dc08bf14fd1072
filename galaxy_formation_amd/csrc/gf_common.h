// Shared definitions of the gfx950 force-path kernels (wave64, HIP).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Tree constants of the reference octree (cpp/nbody.cpp:33-35).
#define GF_LEAF_SIZE 8
#define GF_GROUP_SIZE 64
#define GF_MAX_DEPTH 48
// Octant digits held by one 63-bit Morton key (3 bits per level).
#define GF_KEY_LEVELS 21

#define GF_WAVE 64

// ------------------------------------------------------------ error state
// Host-side error reporting: every entry point returns 0 on success and a
// negative code on failure, with the message kept for gf_last_error().
void gf_set_error(const char* fmt, ...);

#define GF_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            gf_set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,               \
                         hipGetErrorString(e_));                                          \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

#define GF_LAUNCH_CHECK() GF_HIP(hipGetLastError())

// ------------------------------------------------------------- device helpers
__device__ __forceinline__ int gf_lane() { return __lane_id(); }

__device__ __forceinline__ uint64_t gf_lanemask_lt() {
    const int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive prefix sum over the 64 lanes of a wave, by DPP moves (no LDS):
// row_shr 1, 2, 4, 8 scan each row of 16 lanes (a source before the row
// start reads 0), then row_bcast:15 adds lane 15's total to row 1 and lane
// 47's to row 3, and row_bcast:31 adds lane 31's (rows 0-1) to rows 2 and 3.
// Every lane of the wave must be active.
__device__ __forceinline__ int gf_wave_incl_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

__device__ __forceinline__ int gf_wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// A 64-bit value rotated within each row of 16 lanes by a DPP row_ror.
template <int CTRL>
__device__ __forceinline__ double gf_dpp_ror_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double gf_readlane_d(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Minimum and maximum over the 64 lanes (every lane gets the result), by
// DPP: row_ror 8, 4, 2, 1 leave each lane with its row's extreme, and the
// four rows are combined from lanes 0, 16, 32, 48.  The order of the
// comparisons does not change the result.  Every lane must be active.
__device__ __forceinline__ double gf_wave_min_d(double v) {
    v = fmin(v, gf_dpp_ror_d<0x128>(v));
    v = fmin(v, gf_dpp_ror_d<0x124>(v));
    v = fmin(v, gf_dpp_ror_d<0x122>(v));
    v = fmin(v, gf_dpp_ror_d<0x121>(v));
    return fmin(fmin(gf_readlane_d(v, 0), gf_readlane_d(v, 16)),
                fmin(gf_readlane_d(v, 32), gf_readlane_d(v, 48)));
}

__device__ __forceinline__ double gf_wave_max_d(double v) {
    v = fmax(v, gf_dpp_ror_d<0x128>(v));
    v = fmax(v, gf_dpp_ror_d<0x124>(v));
    v = fmax(v, gf_dpp_ror_d<0x122>(v));
    v = fmax(v, gf_dpp_ror_d<0x121>(v));
    return fmax(fmax(gf_readlane_d(v, 0), gf_readlane_d(v, 16)),
                fmax(gf_readlane_d(v, 32), gf_readlane_d(v, 48)));
}

__device__ __forceinline__ double gf_wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float gf_wave_sum_f(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Octant digit of a key at depth d (1..GF_KEY_LEVELS); digit 1 is the root split.
__device__ __forceinline__ int gf_key_digit(uint64_t key, int d) {
    return int((key >> (3 * (GF_KEY_LEVELS - d))) & 7u);
}

// Key prefix made of the first d digits (d = 0 gives 0 for every key).
__device__ __forceinline__ uint64_t gf_key_prefix(uint64_t key, int d) {
    return d == 0 ? 0ull : (key >> (3 * (GF_KEY_LEVELS - d)));
}

static inline int gf_div_up(long long a, long long b) { return int((a + b - 1) / b); }

// The workgroups of a launch go to the chip's 8 XCDs round-robin (block b on
// XCD b % 8), and every XCD has an L2 of its own.  This logical block index
// gives XCD x one contiguous range of the blocks instead (a bijection of
// 0 .. gridDim.x - 1), so that work neighbouring in memory shares an L2.
#define GF_XCDS 8
__device__ __forceinline__ int gf_xcd_block() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int x = b % GF_XCDS, i = b / GF_XCDS;
    const int per = nb / GF_XCDS, rem = nb % GF_XCDS;
    return x * per + min(x, rem) + i;
}

// Capacities that tests lower to drive the fallback paths (0 = default):
// the walk's LDS stack entries, the neighbour solve's candidate list and
// the nodes above the bottom-up subtrees finished from LDS.
struct GfTestLimits {
    int walk_stack = 0;
    int knn_list = 0;
    int moment_top = 0;
};
GfTestLimits& gf_test_limits();
