// The neighbour cache behind sph.NeighbourFinder (the contract of
// galaxy_formation/sph.py:138-170), held on the GPU between calls.
//
// A full search keeps, per gas particle, its m nearest gas particles as
// global particle indices (cand, rows x m int32), and for every particle of
// the simulation the row it owns (owner, n_total int32, -1 = was not gas).
// A re-selection for the current gas list g maps each candidate back to its
// current local index through a scatter table (loc, n_total int32, -1 =
// not gas now), drops candidates that stopped being gas, and ranks the rest
// by (d^2, local index) to keep the k nearest.  A row that no longer owns a
// cache row, or has fewer than k candidates left, makes the whole call
// report failure, and the caller runs a full search (sph.py:155-166).
//
// Ranking: one wave per row, candidate c on lane c (m <= 64).  Each lane
// counts the candidates that order before its own by broadcasting every
// candidate's key in turn: m wave-wide steps, every lane busy, no sort.
#include "gf_verlet.h"

namespace {

constexpr int kRowsPerBlock = 4;          // waves per block
constexpr int kTB = 256;

__global__ __launch_bounds__(kTB) void k_fill_i32(int n, int v, int* __restrict__ a) {
    for (int i = blockIdx.x * kTB + threadIdx.x; i < n; i += gridDim.x * kTB) a[i] = v;
}

// owner[g[i]] = i; cand[i][c] = g[nb[i][c]] (nb: local indices, n x m)
__global__ __launch_bounds__(kTB) void k_store(int n, int n_total, int m,
                                               const int64_t* __restrict__ g,
                                               const int64_t* __restrict__ nb,
                                               int* __restrict__ owner, int* __restrict__ cand,
                                               int* __restrict__ err) {
    const size_t t = size_t(blockIdx.x) * kTB + threadIdx.x;
    if (t >= size_t(n) * m) return;
    const int i = int(t / m);
    const int64_t local = nb[t];
    if (local < 0 || local >= n || g[local] < 0 || g[local] >= n_total || g[i] < 0 ||
        g[i] >= n_total) {
        atomicOr(err, 1);
        return;
    }
    cand[t] = int(g[local]);
    if (t % m == 0) owner[g[i]] = i;
}

// loc[g[i]] = i (or -1 to undo); row[i] = owner[g[i]]; any gas particle
// without a cache row clears ok
__global__ __launch_bounds__(kTB) void k_scatter(int n, int n_total, const int64_t* __restrict__ g,
                                                 const int* __restrict__ owner, int undo,
                                                 int* __restrict__ loc, int* __restrict__ row,
                                                 int* __restrict__ flags) {
    const int i = blockIdx.x * kTB + threadIdx.x;
    if (i >= n) return;
    const int64_t gi = g[i];
    if (gi < 0 || gi >= n_total) {
        if (!undo) row[i] = -1;            // k_reselect skips the row
        atomicOr(&flags[1], 1);
        return;
    }
    if (undo) {
        loc[gi] = -1;
        return;
    }
    loc[gi] = i;
    const int r = owner[gi];
    row[i] = r;
    if (r < 0) atomicAnd(&flags[0], 0);
}

__device__ __forceinline__ bool key_before(double da, int ia, int ca, double db, int ib, int cb) {
    return da < db || (da == db && (ia < ib || (ia == ib && ca < cb)));
}

__global__ __launch_bounds__(kRowsPerBlock * 64) void k_reselect(
    int n, int m, int k, const double* __restrict__ pos, const int* __restrict__ row,
    const int* __restrict__ cand, const int* __restrict__ loc, int64_t* __restrict__ nb,
    double* __restrict__ dist, int* __restrict__ flags) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    if (i >= n) return;                    // whole waves leave together
    const int r = row[i];
    if (r < 0) return;                     // reported by k_scatter
    int j = -1;
    if (lane < m) j = loc[cand[size_t(r) * m + lane]];   // -1: not gas any more
    double d2 = INFINITY;
    if (j >= 0) {
        const double dx = pos[3 * size_t(i)] - pos[3 * size_t(j)];
        const double dy = pos[3 * size_t(i) + 1] - pos[3 * size_t(j) + 1];
        const double dz = pos[3 * size_t(i) + 2] - pos[3 * size_t(j) + 2];
        d2 = dx * dx + dy * dy + dz * dz;
    }
    const int valid = __popcll(__ballot(j >= 0));
    if (valid < k) {
        if (lane == 0) atomicAnd(&flags[0], 0);
        return;
    }
    // invalid lanes order last: key (inf, INT_MAX)
    const int jk = j >= 0 ? j : 0x7fffffff;
    int rank = 0;
    for (int c = 0; c < m; ++c) {
        const double dc = __shfl(d2, c, 64);
        const int jc = __shfl(jk, c, 64);
        rank += key_before(dc, jc, c, d2, jk, lane);
    }
    if (j >= 0 && rank < k) {
        nb[size_t(i) * k + rank] = j;
        dist[size_t(i) * k + rank] = sqrt(d2);
    }
}

// dst (n x k) = the first k columns of src (n x m), one thread per entry
__global__ __launch_bounds__(kTB) void k_first_columns(int n, int m, int k,
                                                       const int64_t* __restrict__ src_nb,
                                                       const double* __restrict__ src_dist,
                                                       int64_t* __restrict__ nb,
                                                       double* __restrict__ dist) {
    const size_t t = size_t(blockIdx.x) * kTB + threadIdx.x;
    if (t >= size_t(n) * k) return;
    const size_t row = t / k, c = t % k;
    nb[t] = src_nb[row * m + c];
    dist[t] = src_dist[row * m + c];
}

template <typename T>
int reserve(T*& p, size_t& cap, size_t count) {
    if (count <= cap && p) return 0;
    if (p) GF_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    GF_HIP(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (count ? count : 1)));
    cap = count;
    return 0;
}

}  // namespace

int gf_verlet_store(GfVerlet& V, int n_total, int n, int m, const int64_t* g, const int64_t* nb,
                    hipStream_t s) {
    if (m < 1 || m > GF_VERLET_MAX_CAND || n < 0 || n > n_total ||
        size_t(n) * m > size_t(INT32_MAX)) {
        gf_set_error("neighbour cache: %d candidates for %d of %d particles (at most %d "
                     "candidates)", m, n, n_total, GF_VERLET_MAX_CAND);
        return -1;
    }
    if (reserve(V.owner, V.cap_owner, size_t(n_total)) ||
        reserve(V.loc, V.cap_loc, size_t(n_total)) || reserve(V.cand, V.cap_cand, size_t(n) * m) ||
        reserve(V.flags, V.cap_flags, 4))
        return -1;
    const int fill_blocks = std::max(1, std::min(4096, gf_div_up(std::max(n_total, 1), kTB)));
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_blocks), dim3(kTB), 0, s, n_total, -1, V.owner);
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_blocks), dim3(kTB), 0, s, n_total, -1, V.loc);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kTB), 0, s, 4, 0, V.flags);
    if (n > 0)
        hipLaunchKernelGGL(k_store, dim3(gf_div_up(int(size_t(n) * m), kTB)), dim3(kTB), 0, s, n,
                           n_total, m, g, nb, V.owner, V.cand, V.flags + 2);
    GF_LAUNCH_CHECK();
    int err = 0;
    GF_HIP(hipMemcpyAsync(&err, V.flags + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    if (err) {
        gf_set_error("neighbour cache: a neighbour index outside the gas list");
        return -1;
    }
    V.n_total = n_total;
    V.rows = n;
    V.m = m;
    return 0;
}

int gf_verlet_reselect(GfVerlet& V, int n_total, int n, int k, const int64_t* g,
                       const double* pos, int* row_scratch, int64_t* nb, double* dist,
                       hipStream_t s) {
    if (V.rows <= 0 || V.n_total != n_total || k < 1 || k > V.m) return 0;
    if (n <= 0) return 1;
    const int one[2] = {1, 0};
    GF_HIP(hipMemcpyAsync(V.flags, one, sizeof(one), hipMemcpyHostToDevice, s));
    const int nb_scatter = gf_div_up(n, kTB);
    hipLaunchKernelGGL(k_scatter, dim3(nb_scatter), dim3(kTB), 0, s, n, n_total, g, V.owner, 0,
                       V.loc, row_scratch, V.flags);
    hipLaunchKernelGGL(k_reselect, dim3(gf_div_up(n, kRowsPerBlock)), dim3(kRowsPerBlock * 64), 0,
                       s, n, V.m, k, pos, row_scratch, V.cand, V.loc, nb, dist, V.flags);
    hipLaunchKernelGGL(k_scatter, dim3(nb_scatter), dim3(kTB), 0, s, n, n_total, g, V.owner, 1,
                       V.loc, row_scratch, V.flags);
    GF_LAUNCH_CHECK();
    int flags[2] = {0, 0};
    GF_HIP(hipMemcpyAsync(flags, V.flags, sizeof(flags), hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    if (flags[1]) {
        gf_set_error("neighbour cache: a gas index outside [0, n_total)");
        return -1;
    }
    return flags[0];
}

int gf_verlet_export(const GfVerlet& V, int64_t* owner, int64_t* cand, hipStream_t s) {
    if (V.rows <= 0) return 0;
    std::vector<int> o(size_t(V.n_total)), c(size_t(V.rows) * V.m);
    GF_HIP(hipMemcpyAsync(o.data(), V.owner, sizeof(int) * o.size(), hipMemcpyDeviceToHost, s));
    GF_HIP(hipMemcpyAsync(c.data(), V.cand, sizeof(int) * c.size(), hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    if (owner)
        for (size_t i = 0; i < o.size(); ++i) owner[i] = o[i];
    if (cand)
        for (size_t i = 0; i < c.size(); ++i) cand[i] = c[i];
    return V.rows;
}

int gf_first_columns(int n, int m, int k, const int64_t* src_nb, const double* src_dist,
                     int64_t* nb, double* dist, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_first_columns, dim3(gf_div_up(int(size_t(n) * k), kTB)), dim3(kTB), 0, s,
                       n, m, k, src_nb, src_dist, nb, dist);
    GF_LAUNCH_CHECK();
    return 0;
}

void gf_verlet_free(GfVerlet& V) {
    if (V.owner) (void)hipFree(V.owner);
    if (V.loc) (void)hipFree(V.loc);
    if (V.cand) (void)hipFree(V.cand);
    if (V.flags) (void)hipFree(V.flags);
    V = GfVerlet();
}
