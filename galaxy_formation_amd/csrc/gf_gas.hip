// Gas physics beside the force path, on the gas rows of one GPU:
//   * radiative cooling of a primordial gas (galaxy_formation/cooling.py:24-43)
//     and the explicit thermodynamics of a sub-step (simulation.py:97-104);
//   * the star-formation candidates (star_formation.py:37-44): dense, cold
//     gas and its conversion probability;
//   * the supernova feedback (star_formation.py:50-65): the fb_ngb nearest
//     gas particles of each new star (an exact k-nearest query against a
//     tree of the gas), then the heating, applied in the order np.add.at
//     applies the flattened (star, neighbour) list, so a particle heated by
//     several stars gets the same FP64 sum.
// All elementwise work is one pass over HBM; the arithmetic is FP64 with the
// reference's operation order (no FMA contraction).
#include <cmath>

#include "gf_gas.h"
#include "gf_prim.h"

namespace {

constexpr int TB = 256;
constexpr double T_FLOOR = 1e4;     // cooling.py:21
constexpr double T_SF_MAX = 3e4;    // star_formation.py:31

struct Units {
    double rho_cgs, x_h, m_h, myr, sec_per_myr, u_cgs, msun, msun_g;
};

Units units_arg() {
    const GfGasUnits& g = gf_gas_units();
    return Units{g.rho_cgs, g.x_h, g.m_h, g.myr, g.sec_per_myr, g.u_cgs, g.msun, g.msun_g};
}

// Lambda(T) of cooling.py:24-32: Cen (1992) collisional excitation of H and
// He+ with the (1 + sqrt(T / 1e5))^-1 factor, plus bremsstrahlung with a
// Gaunt factor 1.1; zero at or below T_FLOOR.
__device__ __forceinline__ double cooling_lambda(double T) {
#pragma clang fp contract(off)
    const double Tc = fmax(T, 1.0);
    const double f = 1.0 / (1.0 + sqrt(Tc / 1e5));
    const double lam = 7.5e-19 * exp(-118348.0 / Tc) * f +
                       5.54e-17 * pow(Tc, -0.397) * exp(-473638.0 / Tc) * f +
                       (1.42e-27 * 1.1) * sqrt(Tc);
    return Tc > T_FLOOR ? lam : 0.0;
}

// Energy radiated in dt (code units) at density rho (cooling.py:38-41).
__device__ __forceinline__ double cooling_du(double T, double rho, double dt, const Units& c) {
#pragma clang fp contract(off)
    const double rho_cgs = rho * c.rho_cgs;
    const double nH = c.x_h * rho_cgs / c.m_h;
    const double dudt_cgs = cooling_lambda(T) * (nH * nH) / rho_cgs;
    return dudt_cgs * dt * c.myr * c.sec_per_myr / c.u_cgs;
}

__global__ __launch_bounds__(TB) void k_cooling_function(int n, const double* __restrict__ T,
                                                         double* __restrict__ lam) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i < n) lam[i] = cooling_lambda(T[i]);
}

__global__ __launch_bounds__(TB) void k_cool(int n, const double* __restrict__ u,
                                             const double* __restrict__ T,
                                             const double* __restrict__ rho, double dt,
                                             const uint8_t* __restrict__ blocked, double u_floor,
                                             Units c, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const double du = (blocked && blocked[i]) ? 0.0 : cooling_du(T[i], rho[i], dt, c);
    out[i] = fmax(u[i] - du, u_floor);
}

// One gas row of simulation.py:102-104.  Rows whose index is outside
// [0, n_total) are skipped and counted in err.
__global__ __launch_bounds__(TB) void k_thermo(int ng, const int64_t* __restrict__ g, int n_total,
                                               double* __restrict__ u,
                                               const double* __restrict__ dudt,
                                               const double* __restrict__ rho,
                                               const double* __restrict__ ncu, double t,
                                               double dt, double u_per_K, double u_floor,
                                               Units c, int* __restrict__ err) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= ng) return;
    const int64_t j = g ? g[i] : i;
    if (j < 0 || j >= n_total) {
        atomicAdd(err, 1);
        return;
    }
    const double u1 = fmax(u[j] + dt * dudt[i], u_floor);
    const bool blocked = ncu ? ncu[j] > t : false;
    const double du = blocked ? 0.0 : cooling_du(u1 / u_per_K, rho[i], dt, c);
    u[j] = fmax(u1 - du, u_floor);
}

// Dense, cold rows (star_formation.py:39): flag[i] = 1.
__global__ __launch_bounds__(TB) void k_sf_flag(int ng, const int64_t* __restrict__ g, int n_total,
                                                const double* __restrict__ u,
                                                const double* __restrict__ rho, double u_per_K,
                                                double rho_sf, int* __restrict__ flag,
                                                int* __restrict__ err) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= ng) return;
    const int64_t j = g ? g[i] : i;
    int f = 0;
    if (j < 0 || j >= n_total) {
        atomicAdd(err, 1);
    } else {
        f = (rho[i] > rho_sf) && (u[j] / u_per_K < T_SF_MAX);
    }
    flag[i] = f;
}

// Rows in order and p = 1 - exp(-sf_eff dt / t_ff), t_ff = sqrt(3 pi / (32 rho))
// (star_formation.py:42-43).
__global__ __launch_bounds__(TB) void k_sf_compact(int ng, const int* __restrict__ off,
                                                   const double* __restrict__ rho,
                                                   double neg_eff_dt, int64_t* __restrict__ rows,
                                                   double* __restrict__ p) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= ng || off[i + 1] == off[i]) return;
    const int o = off[i];
    const double t_ff = sqrt(3.0 * M_PI / (32.0 * rho[i]));
    rows[o] = i;
    p[o] = 1.0 - exp(neg_eff_dt / t_ff);
}

__global__ __launch_bounds__(TB) void k_sf_draw(int m, const double* __restrict__ p,
                                                const double* __restrict__ uniform,
                                                int* __restrict__ flag) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i < m) flag[i] = uniform[i] < p[i];
}

__global__ __launch_bounds__(TB) void k_sf_pick(int m, const int* __restrict__ off,
                                                const int64_t* __restrict__ rows,
                                                int64_t* __restrict__ out) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i < m && off[i + 1] != off[i]) out[off[i]] = rows[i];
}

// ------------------------------------------------------ k-nearest query
// One wave per query point.  The first radius would hold about 2k sources at
// the mean density of the deepest tree cell on the query's octant path that
// holds >= k sources, capped by the radius that covers that whole cell
// (which holds >= k).  The wave gathers the sources within the radius into
// LDS; fewer than k grows the radius, more than QCAP shrinks it (bracketed
// bisection); then k rounds of a wave-wide arg-min by (d^2, input index)
// give the rows in ascending order.
constexpr int QCAP = 2048;
constexpr int QSCAP = 1024;
constexpr int QUERY_BLOCKS = 4096;
constexpr int QUERY_MAX_ITER = 200;

struct QueryArgs {
    const double4* box;
    const int8_t* leaf;
    const int* start;
    const int* count;
    const int* child8;
    const double* tx;
    const double* ty;
    const double* tz;
    const uint32_t* idx;
    const double* q;
    int nq, k;
    int64_t* nb;
    double* dist;
    int* state;
};

// Pops per step from a LIFO of capacity cap holding sn entries: a batch of 64
// while their children fit with room to spare, else one (depth first needs
// at most 7 more entries per tree level).
__device__ __forceinline__ int q_stack_take(int sn, int cap) {
    return (cap - sn >= 7 * 64 + 8 * (GF_KEY_LEVELS + 1)) ? min(64, sn) : 1;
}

__device__ __forceinline__ bool q_box_meets(const double4& b, double x, double y, double z,
                                            double r2) {
#pragma clang fp contract(off)
    const double ex = fmax(0.0, fabs(x - b.x) - b.w);
    const double ey = fmax(0.0, fabs(y - b.y) - b.w);
    const double ez = fmax(0.0, fabs(z - b.z) - b.w);
    return ex * ex + ey * ey + ez * ez <= r2;
}

// Sources within r2 of (qx, qy, qz): their (d^2, input index) go to the list
// (the first QCAP of them); returns how many there are.
__device__ int q_gather(const QueryArgs& a, double qx, double qy, double qz, double r2, int* stack,
                        double* ld2, int* lid) {
#pragma clang fp contract(off)
    const int lane = __lane_id();
    const uint64_t lt = gf_lanemask_lt();
    int sn = 1, cnt = 0;
    if (lane == 0) stack[0] = 0;
    __builtin_amdgcn_wave_barrier();
    while (sn > 0) {
        const int take = q_stack_take(sn, QSCAP);
        sn -= take;
        const int node = lane < take ? stack[sn + lane] : -1;
        __builtin_amdgcn_wave_barrier();
        int push = 0, first = 0, nent = 0;
        int4 c_lo = make_int4(-1, -1, -1, -1), c_hi = c_lo;
        if (node >= 0 && q_box_meets(a.box[node], qx, qy, qz, r2)) {
            if (a.leaf[node]) {
                first = a.start[node];
                nent = a.count[node];
            } else {
                const int4* cs = reinterpret_cast<const int4*>(a.child8 + 8 * size_t(node));
                c_lo = cs[0];
                c_hi = cs[1];
                push = (c_lo.x >= 0) + (c_lo.y >= 0) + (c_lo.z >= 0) + (c_lo.w >= 0) +
                       (c_hi.x >= 0) + (c_hi.y >= 0) + (c_hi.z >= 0) + (c_hi.w >= 0);
            }
        }
        const int pinc = gf_wave_incl_scan(push);
        {
            int o = sn + pinc - push;
            const int cc[8] = {c_lo.x, c_lo.y, c_lo.z, c_lo.w, c_hi.x, c_hi.y, c_hi.z, c_hi.w};
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (cc[k] >= 0) stack[o++] = cc[k];
        }
        sn += __builtin_amdgcn_readlane(pinc, 63);
        // the opened leaves' sources, flattened over the lanes
        const int qinc = gf_wave_incl_scan(nent);
        const int tot = __builtin_amdgcn_readlane(qinc, 63);
        for (int e0 = 0; e0 < tot; e0 += 64) {
            const int e = e0 + lane;
            int own = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1)
                if (__shfl(qinc, own + step - 1, 64) <= e) own += step;
            const int ofirst = __shfl(first, own, 64);
            const int obase = __shfl(qinc - nent, own, 64);
            bool in = false;
            double d2 = 0;
            int id = 0;
            if (e < tot) {
                const int j = ofirst + (e - obase);
                const double dx = qx - a.tx[j], dy = qy - a.ty[j], dz = qz - a.tz[j];
                d2 = dx * dx + dy * dy + dz * dz;
                in = d2 <= r2;
                id = int(a.idx[j]);
            }
            const uint64_t m = __ballot(in);
            if (in) {
                const int slot = cnt + __popcll(m & lt);
                if (slot < QCAP) {
                    ld2[slot] = d2;
                    lid[slot] = id;
                }
            }
            cnt += __popcll(m);
        }
        __builtin_amdgcn_wave_barrier();
    }
    return cnt;
}

__global__ __launch_bounds__(64) void k_knn_query(QueryArgs a) {
#pragma clang fp contract(off)
    __shared__ double ld2[QCAP];
    __shared__ int lid[QCAP];
    __shared__ int stack[QSCAP];
    const int lane = __lane_id();
    for (int q = blockIdx.x; q < a.nq; q += gridDim.x) {
        const double qx = a.q[3 * size_t(q)], qy = a.q[3 * size_t(q) + 1], qz = a.q[3 * size_t(q) + 2];
        // deepest cell on the octant path holding >= k sources
        int node = 0;
        while (!a.leaf[node]) {
            const double4 b = a.box[node];
            const int o = int(qx > b.x) | (int(qy > b.y) << 1) | (int(qz > b.z) << 2);
            const int c = a.child8[8 * size_t(node) + o];
            if (c < 0 || a.count[c] < a.k) break;
            node = c;
        }
        const double4 b = a.box[node];
        const double fx = fabs(qx - b.x) + b.w, fy = fabs(qy - b.y) + b.w, fz = fabs(qz - b.z) + b.w;
        // the radius that holds about 2k sources at the cell's mean density,
        // at most the one that covers the whole cell (which holds >= k)
        const double side = 2.0 * b.w;
        const double dens = double(a.count[node]) / (side * side * side);
        const double r_est = cbrt(6.0 * a.k / (4.0 * M_PI * dens));
        double r = fmin(sqrt(fx * fx + fy * fy + fz * fz) * (1.0 + 1e-12), r_est);
        double lo = 0.0, hi = INFINITY;
        int cnt = 0, it = 0;
        while (true) {
            cnt = q_gather(a, qx, qy, qz, r * r, stack, ld2, lid);
            if (cnt >= a.k && cnt <= QCAP) break;
            if (++it > QUERY_MAX_ITER) break;
            if (cnt < a.k) {
                lo = r;
                r = hi < INFINITY ? 0.5 * (lo + hi) : 2.0 * r;
            } else {
                hi = r;
                const double rn = r * cbrt(0.5 * QCAP / double(cnt));
                r = rn > lo ? rn : 0.5 * (lo + hi);
            }
        }
        if (cnt < a.k || cnt > QCAP) {
            if (lane == 0) atomicAdd(&a.state[0], 1);
            continue;
        }
        // k rounds of arg-min by (d^2, input index)
        for (int s = 0; s < a.k; ++s) {
            double bd = INFINITY;
            int bi = 0x7fffffff, bs = -1;
            for (int c = lane; c < cnt; c += 64) {
                const double d = ld2[c];
                const int i = lid[c];
                if (d < bd || (d == bd && i < bi)) {
                    bd = d;
                    bi = i;
                    bs = c;
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const double od = __shfl_xor(bd, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                const int os = __shfl_xor(bs, o, 64);
                if (od < bd || (od == bd && oi < bi)) {
                    bd = od;
                    bi = oi;
                    bs = os;
                }
            }
            if (lane == 0) {
                a.nb[size_t(q) * a.k + s] = bi;
                if (a.dist) a.dist[size_t(q) * a.k + s] = sqrt(bd);
                ld2[bs] = INFINITY;
                lid[bs] = 0x7fffffff;
            }
            __builtin_amdgcn_wave_barrier();
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------ feedback heating
// du of each (star, neighbour) entry, star_formation.py:61-63:
//   E = e_per_g m_star MSUN MSUN_G,  du = E / k / (m_target MSUN MSUN_G) / U_CGS,
// and the sort pairs (target, entry).  Targets outside [0, n_total) are
// counted in err and get a key past every particle.
__global__ __launch_bounds__(TB) void k_fb_prep(int m, int k, const int64_t* __restrict__ targets,
                                                const double* __restrict__ new_mass,
                                                double e_per_g, int n_total,
                                                const double* __restrict__ mass, Units c,
                                                uint64_t* __restrict__ key,
                                                uint32_t* __restrict__ val,
                                                double* __restrict__ du, int* __restrict__ err) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= m) return;
    const int64_t t = targets[i];
    val[i] = uint32_t(i);
    if (t < 0 || t >= n_total) {
        atomicAdd(err, 1);
        key[i] = uint64_t(n_total);
        du[i] = 0.0;
        return;
    }
    key[i] = uint64_t(t);
    const double E = e_per_g * new_mass[i / k] * c.msun * c.msun_g;
    du[i] = E / double(k) / (mass[t] * c.msun * c.msun_g) / c.u_cgs;
}

// One thread per run of equal targets (sorted, entries ascending inside a
// run): u[t] += du in entry order, as np.add.at does.
__global__ __launch_bounds__(TB) void k_fb_apply(int m, int n_total,
                                                 const uint64_t* __restrict__ key,
                                                 const uint32_t* __restrict__ val,
                                                 const double* __restrict__ du,
                                                 double* __restrict__ u,
                                                 double* __restrict__ ncu, double until) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= m) return;
    const uint64_t t = key[i];
    if (t >= uint64_t(n_total) || (i > 0 && key[i - 1] == t)) return;
    double v = u[t];
    for (int j = i; j < m && key[j] == t; ++j) v = v + du[val[j]];
    u[t] = v;
    ncu[t] = until;
}

template <typename T>
int grow(T*& p, size_t& cap, size_t n) {
    if (n <= cap && p) return 0;
    if (p) GF_HIP(hipFree(p));
    p = nullptr;
    GF_HIP(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (n ? n : 1)));
    cap = n;
    return 0;
}

int ensure_state(GfGas& G) {
    if (!G.state) {
        GF_HIP(hipMalloc(&G.state, 8 * sizeof(int)));
        GF_HIP(hipMemset(G.state, 0, 8 * sizeof(int)));
    }
    return 0;
}

// Scan scratch for n flags: flag (n + 1) and the scan partials.
int ensure_flags(GfGas& G, int n) {
    if (grow(G.flag, G.cap_flag, size_t(n) + 1)) return -1;
    const size_t need = gf_scan_temp_ints(n + 1);
    if (need <= G.cap_scan && G.scan_tmp) return 0;
    // a new scan scratch starts with its look-back state cleared
    if (grow(G.scan_tmp, G.cap_scan, need)) return -1;
    return gf_scan_temp_clear(G.scan_tmp, n + 1);
}

// Reads and clears the error counter at state[2].
int take_errors(GfGas& G, hipStream_t s, int* out) {
    int e = 0;
    GF_HIP(hipMemcpyAsync(&e, G.state + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    if (e) GF_HIP(hipMemsetAsync(G.state + 2, 0, sizeof(int), s));
    *out = e;
    return 0;
}

}  // namespace

const GfGasUnits& gf_gas_units() {
    static const GfGasUnits u = [] {
        GfGasUnits v;
        const double g_kpc = 4.30091e-6, kpc = 3.5, kpc_cm = 3.086e21, yr_s = 3.156e7;
        v.msun = 4e10;
        v.msun_g = 1.989e33;
        v.m_h = 1.6726e-24;
        v.x_h = 0.76;
        const double kms = std::sqrt(g_kpc * v.msun / kpc);
        v.myr = kpc / kms * 977.79;
        v.rho_cgs = v.msun * v.msun_g / std::pow(kpc * kpc_cm, 3.0);
        v.u_cgs = (kms * 1e5) * (kms * 1e5);
        v.sec_per_myr = 1e6 * yr_s;
        return v;
    }();
    return u;
}

void gf_gas_free(GfGas& G) {
    gf_tree_free(G.tree);
    if (G.state) (void)hipFree(G.state);
    if (G.flag) (void)hipFree(G.flag);
    if (G.scan_tmp) (void)hipFree(G.scan_tmp);
    if (G.key) (void)hipFree(G.key);
    if (G.val) (void)hipFree(G.val);
    if (G.key2) (void)hipFree(G.key2);
    if (G.val2) (void)hipFree(G.val2);
    if (G.du) (void)hipFree(G.du);
    if (G.sort_tmp) (void)hipFree(G.sort_tmp);
    G = GfGas();
}

int gf_cooling_function(int n, const double* T, double* lam, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_cooling_function, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, T, lam);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_cool(int n, const double* u, const double* T, const double* rho, double dt,
            const uint8_t* blocked, double u_floor, double* out, hipStream_t s) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_cool, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, u, T, rho, dt, blocked,
                       u_floor, units_arg(), out);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_thermo_launch(GfGas& G, int ng, const int64_t* g, int n_total, double* u,
                     const double* dudt, const double* rho, const double* ncu, double t,
                     double dt, double u_per_K, double u_floor, hipStream_t s) {
    if (ng <= 0) return 0;
    if (ensure_state(G)) return -1;
    hipLaunchKernelGGL(k_thermo, dim3(gf_div_up(ng, TB)), dim3(TB), 0, s, ng, g, n_total, u, dudt,
                       rho, ncu, t, dt, u_per_K, u_floor, units_arg(), G.state + 2);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_gas_errors(GfGas& G, hipStream_t s) {
    if (!G.state) return 0;
    int e = 0;
    if (take_errors(G, s, &e)) return -1;
    return e;
}

int gf_sf_candidates(GfGas& G, int ng, const int64_t* g, int n_total, const double* u,
                     const double* rho, double u_per_K, double rho_sf, double sf_eff, double dt,
                     int64_t* rows, double* p, hipStream_t s) {
    if (ng <= 0) return 0;
    if (ensure_state(G) || ensure_flags(G, ng)) return -1;
    hipLaunchKernelGGL(k_sf_flag, dim3(gf_div_up(ng, TB)), dim3(TB), 0, s, ng, g, n_total, u, rho,
                       u_per_K, rho_sf, G.flag, G.state + 2);
    GF_LAUNCH_CHECK();
    if (gf_scan_excl(G.flag, G.flag, ng, G.scan_tmp, s)) return -1;
    hipLaunchKernelGGL(k_sf_compact, dim3(gf_div_up(ng, TB)), dim3(TB), 0, s, ng, G.flag, rho,
                       -sf_eff * dt, rows, p);
    GF_LAUNCH_CHECK();
    int count = 0;
    GF_HIP(hipMemcpyAsync(&count, G.flag + ng, sizeof(int), hipMemcpyDeviceToHost, s));
    int e = 0;
    if (take_errors(G, s, &e)) return -1;
    if (e) {
        gf_set_error("star formation: %d gas indices outside [0, %d)", e, n_total);
        return -1;
    }
    return count;
}

int gf_sf_select(GfGas& G, int m, const int64_t* rows, const double* p, const double* uniform,
                 int64_t* out, hipStream_t s) {
    if (m <= 0) return 0;
    if (ensure_flags(G, m)) return -1;
    hipLaunchKernelGGL(k_sf_draw, dim3(gf_div_up(m, TB)), dim3(TB), 0, s, m, p, uniform, G.flag);
    GF_LAUNCH_CHECK();
    if (gf_scan_excl(G.flag, G.flag, m, G.scan_tmp, s)) return -1;
    hipLaunchKernelGGL(k_sf_pick, dim3(gf_div_up(m, TB)), dim3(TB), 0, s, m, G.flag, rows, out);
    GF_LAUNCH_CHECK();
    int count = 0;
    GF_HIP(hipMemcpyAsync(&count, G.flag + m, sizeof(int), hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    return count;
}

int gf_knn_query(GfGas& G, int n_src, const double* src, int n_q, const double* q, int k,
                 int64_t* nb, double* dist, hipStream_t s) {
    if (k < 1 || k > n_src || k > 256) {
        gf_set_error("k-nearest query: k = %d must be in [1, min(n_src = %d, 256)]", k, n_src);
        return -1;
    }
    if (n_q <= 0) return 0;
    if (ensure_state(G)) return -1;
    if (gf_tree_build(G.tree, n_src, src, nullptr, 0.5f, s, /*with_com=*/false)) return -1;
    GF_HIP(hipMemsetAsync(G.state, 0, sizeof(int), s));
    QueryArgs a;
    a.box = G.tree.box; a.leaf = G.tree.leaf; a.start = G.tree.start; a.count = G.tree.count;
    a.child8 = G.tree.child8; a.tx = G.tree.tx; a.ty = G.tree.ty; a.tz = G.tree.tz;
    a.idx = G.tree.idx; a.q = q; a.nq = n_q; a.k = k; a.nb = nb; a.dist = dist; a.state = G.state;
    hipLaunchKernelGGL(k_knn_query, dim3(std::min(n_q, QUERY_BLOCKS)), dim3(64), 0, s, a);
    GF_LAUNCH_CHECK();
    int bad = 0;
    GF_HIP(hipMemcpyAsync(&bad, G.state, sizeof(int), hipMemcpyDeviceToHost, s));
    GF_HIP(hipStreamSynchronize(s));
    if (bad) {
        gf_set_error("k-nearest query: %d query points found no radius holding between k and %d "
                     "sources", bad, QCAP);
        return -1;
    }
    return 0;
}

int gf_feedback_heat(GfGas& G, int n_new, int k, const int64_t* targets, const double* new_mass,
                     double e_per_g, double until, int n_total, const double* mass, double* u,
                     double* no_cool_until, hipStream_t s) {
    if (n_new <= 0) return 0;
    if (k < 1 || (long long)n_new * k > 0x7fffffffLL) {
        gf_set_error("feedback: bad k = %d for %d stars", k, n_new);
        return -1;
    }
    const int m = n_new * k;
    if (ensure_state(G)) return -1;
    if (size_t(m) > G.cap_pair || !G.key) {
        size_t c0 = G.cap_pair, c1 = G.cap_pair, c2 = G.cap_pair, c3 = G.cap_pair, c4 = G.cap_pair;
        if (grow(G.key, c0, size_t(m)) || grow(G.val, c1, size_t(m)) ||
            grow(G.key2, c2, size_t(m)) || grow(G.val2, c3, size_t(m)) ||
            grow(G.du, c4, size_t(m)))
            return -1;
        G.cap_pair = size_t(m);
    }
    if (grow(G.sort_tmp, G.cap_sort, gf_sort_temp_ints(m))) return -1;
    hipLaunchKernelGGL(k_fb_prep, dim3(gf_div_up(m, TB)), dim3(TB), 0, s, m, k, targets, new_mass,
                       e_per_g, n_total, mass, units_arg(), G.key, G.val, G.du, G.state + 2);
    GF_LAUNCH_CHECK();
    int e = 0;
    if (take_errors(G, s, &e)) return -1;
    if (e) {
        gf_set_error("feedback: %d targets outside [0, %d)", e, n_total);
        return -1;
    }
    int bits = 8;
    while (bits < 64 && (uint64_t(n_total) >> bits) != 0) bits += 8;
    int in_alt = 0;
    if (gf_radix_sort_pairs(G.key, G.val, G.key2, G.val2, m, 0, bits, G.sort_tmp, s, &in_alt))
        return -1;
    const uint64_t* ks = in_alt ? G.key2 : G.key;
    const uint32_t* vs = in_alt ? G.val2 : G.val;
    hipLaunchKernelGGL(k_fb_apply, dim3(gf_div_up(m, TB)), dim3(TB), 0, s, m, n_total, ks, vs, G.du,
                       u, no_cool_until, until);
    GF_LAUNCH_CHECK();
    return 0;
}
