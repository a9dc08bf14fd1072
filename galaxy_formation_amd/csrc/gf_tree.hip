// Octree build and refit on the device.
//
// The tree is the one cpp/nbody.cpp:202-256 builds: root box from the bounding
// box (half-width x 1.0001 + 1e-12), octant o = [x > cx] | [y > cy] << 1 |
// [z > cz] << 2 with child centres c +- h from repeated FP64 halving, a cell
// is a leaf when it holds <= 8 particles, nodes in DFS preorder with a `next`
// index, particles in a leaf in ascending input index.
//
// Instead of recursive partitioning, every particle gets a 63-bit key whose
// 21 octant digits come from that same FP64 comparison chain, so the key
// order is the reference's partition order.  A stable radix sort of
// (key, index) gives the tree order.  With L_j the number of leading digits
// keys j-1 and j share, the cell of depth d that starts at position j exists
// iff L_j < d, and it holds more than 8 particles iff the next 8 keys share
// its d digits.  Both facts only need a window of 8 positions, so the number
// of nodes that start at each position, their DFS indices (an exclusive
// scan: preorder is the order of (start, depth)) and `next` (the DFS index
// of the first node starting at the cell's end) are all computed in parallel.
// Centres of mass are summed bottom-up in FP64, children in octant order,
// in two launches: subtrees of at most 256 nodes, each finished by one block
// in LDS, then the nodes above them.
#include "gf_lookback.h"
#include "gf_prim.h"
#include "gf_tree.h"

namespace {

constexpr int TB = 256;

__device__ __forceinline__ int8_t lcp_levels(uint64_t a, uint64_t b) {
    const uint64_t x = a ^ b;
    if (x == 0) return GF_KEY_LEVELS;
    return int8_t((__clzll(x) - 1) / 3);
}

__global__ __launch_bounds__(TB) void k_bbox_partial(const double* __restrict__ pos, int n,
                                                     double* __restrict__ part,
                                                     uint32_t* __restrict__ idx) {
    __shared__ double lds[4][6];
    double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int i = blockIdx.x * TB + threadIdx.x; i < n; i += gridDim.x * TB) {
        idx[i] = uint32_t(i);
        const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        v[0] = fmin(v[0], x); v[1] = fmin(v[1], y); v[2] = fmin(v[2], z);
        v[3] = fmax(v[3], x); v[4] = fmax(v[4], y); v[5] = fmax(v[5], z);
    }
    const int w = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) v[k] = gf_wave_min_d(v[k]);
    for (int k = 3; k < 6; ++k) v[k] = gf_wave_max_d(v[k]);
    if (__lane_id() == 0)
        for (int k = 0; k < 6; ++k) lds[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double r = lds[0][k];
        for (int ww = 1; ww < 4; ++ww) r = k < 3 ? fmin(r, lds[ww][k]) : fmax(r, lds[ww][k]);
        part[6 * blockIdx.x + k] = r;
    }
}

// Root box exactly as nbody.cpp:218-219 (one block reduces the partials).
__global__ __launch_bounds__(TB) void k_bbox_final(const double* __restrict__ part, int nb,
                                                   double* __restrict__ root) {
#pragma clang fp contract(off)
    __shared__ double lds[4][6];
    double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int b = threadIdx.x; b < nb; b += TB) {
        for (int k = 0; k < 3; ++k) v[k] = fmin(v[k], part[6 * b + k]);
        for (int k = 3; k < 6; ++k) v[k] = fmax(v[k], part[6 * b + k]);
    }
    const int w = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) v[k] = gf_wave_min_d(v[k]);
    for (int k = 3; k < 6; ++k) v[k] = gf_wave_max_d(v[k]);
    if (__lane_id() == 0)
        for (int k = 0; k < 6; ++k) lds[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 6; ++k) {
        double r = lds[0][k];
        for (int ww = 1; ww < 4; ++ww) r = k < 3 ? fmin(r, lds[ww][k]) : fmax(r, lds[ww][k]);
        v[k] = r;
    }
    root[0] = 0.5 * (v[0] + v[3]);
    root[1] = 0.5 * (v[1] + v[4]);
    root[2] = 0.5 * (v[2] + v[5]);
    const double ext = fmax(fmax(v[3] - v[0], v[4] - v[1]), v[5] - v[2]);
    root[3] = 0.5 * ext * 1.0001 + 1e-12;
}

// Key digits from the FP64 comparison chain of nbody.cpp:80 and :92-96.
__global__ __launch_bounds__(TB) void k_keys(const double* __restrict__ pos, int n,
                                             const double* __restrict__ root,
                                             uint64_t* __restrict__ key) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    double cx = root[0], cy = root[1], cz = root[2], h = root[3];
    uint64_t k = 0;
#pragma unroll
    for (int t = 0; t < GF_KEY_LEVELS; ++t) {
        const int o = int(x > cx) | (int(y > cy) << 1) | (int(z > cz) << 2);
        k = (k << 3) | uint64_t(o);
        h = 0.5 * h;
        cx = cx + ((o & 1) ? h : -h);
        cy = cy + ((o & 2) ? h : -h);
        cz = cz + ((o & 4) ? h : -h);
    }
    key[i] = k;
}

__global__ __launch_bounds__(TB) void k_lcp(const uint64_t* __restrict__ key, int n,
                                            int8_t* __restrict__ lcp) {
    const int j = blockIdx.x * TB + threadIdx.x;
    if (j >= n) return;
    lcp[j] = j == 0 ? int8_t(-1) : lcp_levels(key[j - 1], key[j]);
}

// Number of nodes that start at sorted position j = i + Q < n (a chain of
// nested cells of depths d0 .. d_last, the last one a leaf; d0 = lcp[j] + 1),
// from w[k] = lcp[i - 8 + k], k in [0, 20): the entries this reads all lie
// in [0, n).
template <int Q>
__device__ __forceinline__ int chain_count(const int (&w)[20], int n, int j) {
    const int L = w[8 + Q];
    const int d0 = L + 1;
    int c = 0;
    bool node = (j == 0);
    if (j > 0 && L < GF_KEY_LEVELS) {
        // the parent cell (depth L) starts at ip = j + ipo, the last position
        // before j with lcp < L; it is internal iff it holds more than 8
        int ipo = 0;
#pragma unroll
        for (int t = 1; t <= GF_LEAF_SIZE; ++t)
            if (ipo == 0 && j - t >= 0 && w[8 + Q - t] < L) ipo = -t;
        if (ipo == 0) {
            node = true;                          // parent spans >= 10 positions
        } else {
            node = j + ipo + GF_LEAF_SIZE <= n - 1;
#pragma unroll
            for (int u = 1; u < GF_LEAF_SIZE; ++u)
                if (node && u <= ipo + GF_LEAF_SIZE) node = w[8 + Q + u] >= L;
        }
    }
    if (node) {
        int m8 = -1;                              // cells (d, j) with d <= m8 hold > 8
        if (j + GF_LEAF_SIZE <= n - 1) {
            m8 = GF_KEY_LEVELS;
#pragma unroll
            for (int u = 1; u <= GF_LEAF_SIZE; ++u) m8 = min(m8, w[8 + Q + u]);
        }
        int dl = m8 >= d0 ? m8 + 1 : d0;
        dl = min(dl, GF_KEY_LEVELS);
        c = dl - d0 + 1;
    }
    return c;
}

// Node bases in one single-pass launch: chain[j] = the nodes that start
// before sorted position j (chain[n] = the node count, also counters[0];
// counters[1] and [2] restart at 0).
struct ChainOp {
    const int8_t* lcp;
    int* chain;
    int8_t* d0o;
    int* counters;
    // a thread's 4 positions read lcp[i - 8, i + 12): five aligned 4-byte
    // loads inside the array, else byte loads of the positions in [0, n)
    __device__ void values(int i, int n, int (&v)[4]) const {
        int w[20];
        if (i >= 8 && i + 12 <= n) {
            const int* src = reinterpret_cast<const int*>(lcp + i - 8);
#pragma unroll
            for (int d = 0; d < 5; ++d) {
                const int x = src[d];
#pragma unroll
                for (int b = 0; b < 4; ++b) w[4 * d + b] = int(int8_t(uint32_t(x) >> (8 * b)));
            }
        } else {
#pragma unroll
            for (int k = 0; k < 20; ++k) {
                const int p = i - 8 + k;
                w[k] = p >= 0 && p < n ? int(lcp[p]) : 0;
            }
        }
        v[0] = i < n ? chain_count<0>(w, n, i) : 0;
        v[1] = i + 1 < n ? chain_count<1>(w, n, i + 1) : 0;
        v[2] = i + 2 < n ? chain_count<2>(w, n, i + 2) : 0;
        v[3] = i + 3 < n ? chain_count<3>(w, n, i + 3) : 0;
        if (i + 3 < n) {
            const uint32_t d = (uint32_t(uint8_t(w[8] + 1))) | (uint32_t(uint8_t(w[9] + 1)) << 8) |
                               (uint32_t(uint8_t(w[10] + 1)) << 16) |
                               (uint32_t(uint8_t(w[11] + 1)) << 24);
            *reinterpret_cast<uint32_t*>(d0o + i) = d;
        } else {
            for (int q = 0; q < 3; ++q)
                if (i + q < n) d0o[i + q] = int8_t(w[8 + q] + 1);
        }
    }
    __device__ void put(int i, int n, const int (&)[4], const int (&o)[4]) const {
        if (i + 3 < n) {
            *reinterpret_cast<int4*>(chain + i) = make_int4(o[0], o[1], o[2], o[3]);
        } else {
            for (int q = 0; q < 3; ++q)
                if (i + q < n) chain[i + q] = o[q];
        }
    }
    __device__ void total(int n, int s) const {
        chain[n] = s;
        counters[0] = s;
        counters[1] = 0;
        counters[2] = 0;
    }
};

__global__ __launch_bounds__(TB) void k_node_init(const int* __restrict__ counters,
                                                  int* __restrict__ child8,
                                                  int* __restrict__ dmeta) {
    const int nn = counters[0];
    for (int i = blockIdx.x * TB + threadIdx.x; i < nn; i += gridDim.x * TB) {
        int4* c = reinterpret_cast<int4*>(child8 + 8 * size_t(i));
        c[0] = make_int4(-1, -1, -1, -1);
        c[1] = make_int4(-1, -1, -1, -1);
    }
    if (blockIdx.x == 0 && threadIdx.x < 32) dmeta[threadIdx.x] = 0;
}

// Writes the nodes that start at position j: range, depth, leaf flag, next,
// parent, geometric box, and registers each in its parent's octant slot.
__global__ __launch_bounds__(TB) void k_emit(
    const uint64_t* __restrict__ key, const int* __restrict__ base, const int8_t* __restrict__ d0a,
    int n, const double* __restrict__ root, int* __restrict__ start, int* __restrict__ count,
    int* __restrict__ next, int* __restrict__ parent, int8_t* __restrict__ leaf,
    int8_t* __restrict__ depth, double4* __restrict__ box, int* __restrict__ child8,
    int* __restrict__ counters) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * TB + threadIdx.x;
    if (j >= n) return;
    const int b = base[j];
    const int c = base[j + 1] - b;
    if (c == 0) return;
    const int d0 = d0a[j];
    const uint64_t kj = key[j];
    double cx = root[0], cy = root[1], cz = root[2], h = root[3];
    for (int t = 1; t < d0; ++t) {
        const int o = gf_key_digit(kj, t);
        h = 0.5 * h;
        cx = cx + ((o & 1) ? h : -h);
        cy = cy + ((o & 2) ? h : -h);
        cz = cz + ((o & 4) ? h : -h);
    }
    int hi_lim = n;                               // cells of one chain are nested
    for (int t = 0; t < c; ++t) {
        const int d = d0 + t;
        if (d >= 1) {
            const int o = gf_key_digit(kj, d);
            h = 0.5 * h;
            cx = cx + ((o & 1) ? h : -h);
            cy = cy + ((o & 2) ? h : -h);
            cz = cz + ((o & 4) ? h : -h);
        }
        const int id = b + t;
        const uint64_t pj = gf_key_prefix(kj, d);
        int lo = j + 1, hi = hi_lim;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (gf_key_prefix(key[mid], d) == pj) lo = mid + 1; else hi = mid;
        }
        const int end = lo;
        hi_lim = end;
        const bool is_leaf = (t == c - 1);
        start[id] = j;
        count[id] = end - j;
        next[id] = base[end];
        depth[id] = int8_t(d);
        leaf[id] = int8_t(is_leaf);
        box[id] = make_double4(cx, cy, cz, h);
        int p;
        if (t > 0) {
            p = id - 1;
        } else if (d == 0) {
            p = -1;
        } else {
            // parent cell: depth d - 1, starts at the first key sharing d - 1 digits
            const uint64_t pp = gf_key_prefix(kj, d - 1);
            int l2 = 0, h2 = j;
            while (l2 < h2) {
                const int mid = (l2 + h2) >> 1;
                if (gf_key_prefix(key[mid], d - 1) < pp) l2 = mid + 1; else h2 = mid;
            }
            p = base[l2] + (d - 1 - d0a[l2]);
        }
        parent[id] = p;
        if (p >= 0) child8[8 * p + gf_key_digit(kj, d)] = id;
        if (is_leaf && end - j > GF_LEAF_SIZE && d < GF_MAX_DEPTH) atomicAdd(&counters[2], 1);
    }
}

// Inside a leaf the reference keeps ascending input index (its partition is
// a stable counting sort that stops at the leaf, nbody.cpp:76-90, :122), while
// the key sort ordered the leaf by deeper digits: sort each leaf by index.
__global__ __launch_bounds__(TB) void k_leaf_sort(const int* __restrict__ counters,
                                                  const int* __restrict__ start,
                                                  const int* __restrict__ count,
                                                  const int8_t* __restrict__ leaf,
                                                  uint32_t* __restrict__ idx,
                                                  uint64_t* __restrict__ key) {
    const int nn = counters[0];
    for (int id = blockIdx.x * TB + threadIdx.x; id < nn; id += gridDim.x * TB) {
        if (!leaf[id]) continue;
        const int b = start[id], e = b + count[id];
        for (int i = b + 1; i < e; ++i) {
            const uint32_t v = idx[i];
            const uint64_t k = key[i];
            int p = i;
            while (p > b && idx[p - 1] > v) {
                idx[p] = idx[p - 1];
                key[p] = key[p - 1];
                --p;
            }
            idx[p] = v;
            key[p] = k;
        }
    }
}

__global__ __launch_bounds__(TB) void k_gather(const double* __restrict__ pos,
                                               const double* __restrict__ mass,
                                               const uint32_t* __restrict__ idx, int n,
                                               const double* __restrict__ xb, double* __restrict__ tx,
                                               double* __restrict__ ty, double* __restrict__ tz,
                                               double* __restrict__ tm,
                                               double* __restrict__ tdisp) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    const int j = int(idx[i]);
    const double x = pos[3 * j], y = pos[3 * j + 1], z = pos[3 * j + 2];
    const double m = mass ? mass[j] : 0.0;      // no masses: a neighbour-search tree
    tx[i] = x; ty[i] = y; tz[i] = z; tm[i] = m;
    if (!tdisp) return;                         // a tree without centres of mass
    double d = 0.0;
    if (xb) {
        const double dx = x - xb[3 * j], dy = y - xb[3 * j + 1], dz = z - xb[3 * j + 2];
        d = m > 0 ? sqrt(dx * dx + dy * dy + dz * dz) : 0.0;
    }
    tdisp[i] = d;
}

// nbody.cpp:98-115 (build) and :297-303 (refit) in FP64: open_r =
// 2 (half + dm) / theta + |com - centre| (dm = 0 right after a build).
__device__ __forceinline__ double4 finalize_node(int id, double M, double sx, double sy,
                                                 double sz, double dm, double theta_d,
                                                 const double4& bx, double4* com,
                                                 double* open_r, double* ndisp) {
#pragma clang fp contract(off)
    const double inv = M > 0 ? 1.0 / M : 0.0;
    const double cx = sx * inv, cy = sy * inv, cz = sz * inv;
    const double dx = cx - bx.x, dy = cy - bx.y, dz = cz - bx.z;
    const double4 c = make_double4(cx, cy, cz, M);
    com[id] = c;
    open_r[id] = 2 * (bx.w + dm) / theta_d + sqrt(dx * dx + dy * dy + dz * dz);
    ndisp[id] = dm;
    return c;
}

// Adds m (3 d d^T - |d|^2 I) to the traceless quadrupole q (xx, yy, zz, xy, xz, yz).
__device__ __forceinline__ void quad_add(double* q, double m, double dx, double dy, double dz) {
#pragma clang fp contract(off)
    const double r2 = dx * dx + dy * dy + dz * dz;
    q[0] += m * (3 * dx * dx - r2);
    q[1] += m * (3 * dy * dy - r2);
    q[2] += m * (3 * dz * dz - r2);
    q[3] += m * (3 * dx * dy);
    q[4] += m * (3 * dx * dz);
    q[5] += m * (3 * dy * dz);
}

// Arrays of the bottom-up pass over one tree.
struct MomentArgs {
    const int* start;
    const int* count;
    const int* next;
    const int* parent;
    const int8_t* leaf;
    const int8_t* depth;
    const int* child8;
    const double4* box;
    const double* tx;
    const double* ty;
    const double* tz;
    const double* tm;
    const double* tdisp;
    double4* com;
    double* open_r;
    double* ndisp;
    double* quad;
    double theta_d;
};

// What a node's moments leave for its parent: the centre of mass and mass,
// the largest displacement and the quadrupole.
struct NodeMoments {
    double4 c;
    double dm;
    double q[6];
};

// The children of an internal node by octant slot: from the tree's arrays
// (node ids), or from the block's LDS window of a subtree (entries, 16 bits
// a slot, 0xFFFF = empty).  An empty slot reads the node itself (entry 0 of
// the window), so the loads of all slots go out together, unconditionally.
struct KidsGlobal {
    static constexpr int QBATCH = 8;     // children's quadrupoles loaded at once
    const MomentArgs& a;
    int self;
    int cs[8];
    __device__ bool has(int o) const { return cs[o] >= 0; }
    __device__ int at(int o) const { return cs[o] >= 0 ? cs[o] : self; }
    __device__ double4 com(int o) const { return a.com[at(o)]; }
    __device__ double nd(int o) const { return a.ndisp[at(o)]; }
    __device__ double q(int o, int j) const { return a.quad[6 * size_t(at(o)) + j]; }
};

struct KidsLds {
    static constexpr int QBATCH = 2;
    const double4* s_com;
    const double* s_nd;
    const double* s_q;
    uint32_t e[4];
    __device__ int entry(int o) const { return int((e[o >> 1] >> (16 * (o & 1))) & 0xFFFFu); }
    __device__ bool has(int o) const { return entry(o) != 0xFFFF; }
    __device__ int at(int o) const { return has(o) ? entry(o) : 0; }
    __device__ double4 com(int o) const { return s_com[at(o)]; }
    __device__ double nd(int o) const { return s_nd[at(o)]; }
    __device__ double q(int o, int j) const { return s_q[6 * at(o) + j]; }
};

// Moments of leaf id (particles [b, b + cnt), box bx) in tree order
// (nbody.cpp:124-131), FP64.  MODE 0: centre of mass, opening radius and
// largest displacement; 1: those and the traceless quadrupole; 2: the
// quadrupole alone, about the held centre of mass `held`.  Results go to the
// tree's arrays and are returned.
template <int MODE>
__device__ NodeMoments leaf_moments(const MomentArgs& a, int id, int b, int cnt,
                                    const double4& bx, const double4& held) {
#pragma clang fp contract(off)
    NodeMoments r;
    const bool small = cnt <= GF_LEAF_SIZE;
    // a leaf of at most GF_LEAF_SIZE particles is loaded in one round (slots
    // past the count read an in-range entry that is not used), then summed
    // in order
    double pm[GF_LEAF_SIZE], px[GF_LEAF_SIZE], py[GF_LEAF_SIZE], pz[GF_LEAF_SIZE];
    double4 c = held;
    r.dm = 0;
    if constexpr (MODE == 2) {
        if (small) {
#pragma unroll
            for (int u = 0; u < GF_LEAF_SIZE; ++u) {
                const int i = b + (u < cnt ? u : 0);
                pm[u] = a.tm[i]; px[u] = a.tx[i]; py[u] = a.ty[i]; pz[u] = a.tz[i];
            }
        }
    } else {
        double M = 0, sx = 0, sy = 0, sz = 0, dm = 0;
        if (small) {
#pragma unroll
            for (int u = 0; u < GF_LEAF_SIZE; ++u) {
                const int i = b + (u < cnt ? u : 0);
                pm[u] = a.tm[i];
                px[u] = a.tx[i];
                py[u] = a.ty[i];
                pz[u] = a.tz[i];
                const double dd = a.tdisp[i];
                if (u < cnt) {
                    M += pm[u];
                    sx += pm[u] * px[u];
                    sy += pm[u] * py[u];
                    sz += pm[u] * pz[u];
                    dm = fmax(dm, dd);
                }
            }
        } else {
            for (int i = b; i < b + cnt; ++i) {
                const double m = a.tm[i];
                M += m;
                sx += m * a.tx[i];
                sy += m * a.ty[i];
                sz += m * a.tz[i];
                dm = fmax(dm, a.tdisp[i]);
            }
        }
        c = finalize_node(id, M, sx, sy, sz, dm, a.theta_d, bx, a.com, a.open_r, a.ndisp);
        r.dm = dm;
    }
    r.c = c;
#pragma unroll
    for (int j = 0; j < 6; ++j) r.q[j] = 0;
    if constexpr (MODE >= 1) {
        double* q = r.q;
        if (small) {
#pragma unroll
            for (int u = 0; u < GF_LEAF_SIZE; ++u)
                if (u < cnt) quad_add(q, pm[u], px[u] - c.x, py[u] - c.y, pz[u] - c.z);
        } else {
            for (int i = b; i < b + cnt; ++i)
                quad_add(q, a.tm[i], a.tx[i] - c.x, a.ty[i] - c.y, a.tz[i] - c.z);
        }
        for (int j = 0; j < 6; ++j) a.quad[6 * size_t(id) + j] = q[j];
    }
    return r;
}

// Moments of internal node id from its children, in octant order
// (nbody.cpp:136-147); the quadrupole sums the children's tensors moved by
// the parallel-axis rule.  Modes as leaf_moments.
template <int MODE, typename Kids>
__device__ NodeMoments internal_moments(const MomentArgs& a, int id, const Kids& k,
                                        const double4& bx, const double4& held) {
#pragma clang fp contract(off)
    NodeMoments r;
    double4 c = held;
    double4 cm[8];
    r.dm = 0;
    if constexpr (MODE != 2) {
        double M = 0, sx = 0, sy = 0, sz = 0, dm = 0;
        double nd[8];
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            cm[o] = k.com(o);
            nd[o] = k.nd(o);
        }
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            if (!k.has(o)) continue;
            M += cm[o].w;
            sx += cm[o].w * cm[o].x;
            sy += cm[o].w * cm[o].y;
            sz += cm[o].w * cm[o].z;
            dm = fmax(dm, nd[o]);
        }
        c = finalize_node(id, M, sx, sy, sz, dm, a.theta_d, bx, a.com, a.open_r, a.ndisp);
        r.dm = dm;
    } else {
#pragma unroll
        for (int o = 0; o < 8; ++o) cm[o] = k.com(o);
    }
    r.c = c;
#pragma unroll
    for (int j = 0; j < 6; ++j) r.q[j] = 0;
    if constexpr (MODE >= 1) {
        // QB children's tensors loaded at a time, summed in octant order
        constexpr int QB = Kids::QBATCH;
        double* q = r.q;
#pragma unroll
        for (int o0 = 0; o0 < 8; o0 += QB) {
            double qq[QB][6];
#pragma unroll
            for (int u = 0; u < QB; ++u)
#pragma unroll
                for (int j = 0; j < 6; ++j) qq[u][j] = k.q(o0 + u, j);
#pragma unroll
            for (int u = 0; u < QB; ++u) {
                const int o = o0 + u;
                if (!k.has(o)) continue;
                for (int j = 0; j < 6; ++j) q[j] += qq[u][j];
                quad_add(q, cm[o].w, cm[o].x - c.x, cm[o].y - c.y, cm[o].z - c.z);
            }
        }
        for (int j = 0; j < 6; ++j) a.quad[6 * size_t(id) + j] = q[j];
    }
    return r;
}

// Bottom-up pass, first half.  The subtree of node id is its DFS range
// [id, next[id]); a subtree root is a node whose subtree has at most
// MOMENT_NCAP nodes while its parent's has more (or the root itself).  The
// block of chunk [c0, c0 + MOMENT_NCAP) finishes the subtrees whose roots lie
// in the chunk: they are disjoint and inside [c0, c0 + 2 MOMENT_NCAP), so one
// LDS window of 2 MOMENT_NCAP entries holds their moments.  A node belongs to
// one of them iff the largest subtree end among the chunk's roots at or
// before it lies past it (a prefix maximum).  Each thread owns entries t and
// t + MOMENT_NCAP of the window and loads what they need first; then the
// leaves (their particles), then the internal nodes one depth at a time,
// deepest first, children read from LDS, so a depth waits on LDS only.
// Nodes whose subtree is larger are listed in top[] (*ntop) for the second
// half.  Three blocks a CU (two with the quadrupoles, whose window takes
// 61 KB of LDS).
constexpr int MOMENT_NCAP = TB;

__device__ __forceinline__ int wave_incl_max(int v) {
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v = max(v, u);
    }
    return v;
}

template <int MODE>
__global__ __launch_bounds__(TB, MODE == 1 ? 2 : 3) void k_moments_subtrees(
    MomentArgs a, const int* __restrict__ counters, int* __restrict__ top, int* __restrict__ ntop) {
    constexpr int W = 2 * MOMENT_NCAP;
    constexpr bool Q = MODE >= 1;
    __shared__ double4 s_com[W];
    __shared__ double s_nd[MODE == 2 ? 1 : W];
    __shared__ double s_q[Q ? 6 * W : 1];
    __shared__ double4 s_box[MODE == 2 ? 1 : W];
    __shared__ int s_wmax[TB / 64];
    __shared__ int s_dlo, s_dhi;
    const int nn = counters[0];
    const int t = threadIdx.x, lane = __lane_id(), w = t >> 6;
    for (int c0 = blockIdx.x * MOMENT_NCAP; c0 < nn; c0 += gridDim.x * MOMENT_NCAP) {
        if (t == 0) {
            s_dlo = 1 << 20;
            s_dhi = -1;
        }
        int endm = 0;
        {
            const int id = c0 + t;
            if (id < nn) {
                const int sz = min(a.next[id], nn) - id;
                if (sz > MOMENT_NCAP) {
                    top[atomicAdd(ntop, 1)] = id;
                } else {
                    const int p = a.parent[id];
                    if (p < 0 || min(a.next[p], nn) - p > MOMENT_NCAP) endm = id + sz;
                }
            }
        }
        int pmax = wave_incl_max(endm);
        if (lane == 63) s_wmax[w] = pmax;
        __syncthreads();
        int all = 0;
        for (int q = 0; q < TB / 64; ++q) {
            if (q < w) pmax = max(pmax, s_wmax[q]);
            all = max(all, s_wmax[q]);
        }
        // what the owned entries need, loaded up front
        bool own[2], leaf[2];
        int dep[2], pb[2], pc[2];
        KidsLds kids[2];
        for (int h = 0; h < 2; ++h) {
            const int id = c0 + t + h * MOMENT_NCAP;
            own[h] = id < nn && (h == 0 ? pmax : all) > id;
            leaf[h] = false;
            dep[h] = pb[h] = pc[h] = 0;
            kids[h] = KidsLds{s_com, s_nd, s_q, {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}};
            if (own[h]) {
                // one round of loads (a leaf's child slots are all empty)
                leaf[h] = a.leaf[id] != 0;
                dep[h] = a.depth[id];
                pb[h] = a.start[id];
                pc[h] = a.count[id];
                if constexpr (MODE != 2) s_box[t + h * MOMENT_NCAP] = a.box[id];
                if constexpr (MODE == 2) s_com[t + h * MOMENT_NCAP] = a.com[id];
                {
                    // the children lie in the same subtree, so inside the window
                    const int4* sl = reinterpret_cast<const int4*>(a.child8 + 8 * size_t(id));
                    const int4 s0 = sl[0], s1 = sl[1];
                    const int cs[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
                    for (int o = 0; o < 8; ++o) {
                        const uint32_t ent = cs[o] >= 0 ? uint32_t(cs[o] - c0) : 0xFFFFu;
                        const uint32_t sh = 16u * uint32_t(o & 1);
                        kids[h].e[o >> 1] = (kids[h].e[o >> 1] & ~(0xFFFFu << sh)) | (ent << sh);
                    }
                }
            }
        }
        __syncthreads();
        int dlo = 1 << 20, dhi = -1;
        for (int h = 0; h < 2; ++h) {
            if (!own[h]) continue;
            const int e = t + h * MOMENT_NCAP;
            if (leaf[h]) {
                const int id = c0 + e;
                const double4 held = MODE == 2 ? s_com[e] : make_double4(0, 0, 0, 0);
                const double4 bx = MODE != 2 ? s_box[e] : make_double4(0, 0, 0, 0);
                const NodeMoments r = leaf_moments<MODE>(a, id, pb[h], pc[h], bx, held);
                if constexpr (MODE != 2) {
                    s_com[e] = r.c;
                    s_nd[e] = r.dm;
                }
                if constexpr (Q)
                    for (int j = 0; j < 6; ++j) s_q[6 * e + j] = r.q[j];
            } else {
                dlo = min(dlo, dep[h]);
                dhi = max(dhi, dep[h]);
            }
        }
        if (dhi >= 0) {
            atomicMin(&s_dlo, dlo);
            atomicMax(&s_dhi, dhi);
        }
        __syncthreads();
        const int d_lo = s_dlo, d_hi = s_dhi;
        for (int d = d_hi; d >= d_lo; --d) {
            for (int h = 0; h < 2; ++h) {
                if (!own[h] || leaf[h] || dep[h] != d) continue;
                const int e = t + h * MOMENT_NCAP;
                const double4 held = MODE == 2 ? s_com[e] : make_double4(0, 0, 0, 0);
                const double4 bx = MODE != 2 ? s_box[e] : make_double4(0, 0, 0, 0);
                const NodeMoments r = internal_moments<MODE>(a, c0 + e, kids[h], bx, held);
                if constexpr (MODE != 2) {
                    s_com[e] = r.c;
                    s_nd[e] = r.dm;
                }
                if constexpr (Q)
                    for (int j = 0; j < 6; ++j) s_q[6 * e + j] = r.q[j];
            }
            __syncthreads();
        }
        // the window is reused by the next chunk
        __syncthreads();
    }
}

// Second half: one block finishes the listed nodes above the subtrees,
// deepest first (their children are subtree roots, done by the first half,
// or listed nodes of the depth below), and clears the list's counter for
// the next pass.  Up to top_max (<= MOMENT_TOP_MAX) listed nodes are sorted
// by depth in LDS (a counting sort) with their children and boxes, so a
// depth is one node a thread and one round of loads (the children's
// moments); a longer list is read from the tree's arrays at every depth.
template <int MODE>
__device__ __forceinline__ NodeMoments top_node(const MomentArgs& a, int id) {
    KidsGlobal kg{a, id, {}};
    const int4* sl = reinterpret_cast<const int4*>(a.child8 + 8 * size_t(id));
    const int4 s0 = sl[0], s1 = sl[1];
    kg.cs[0] = s0.x; kg.cs[1] = s0.y; kg.cs[2] = s0.z; kg.cs[3] = s0.w;
    kg.cs[4] = s1.x; kg.cs[5] = s1.y; kg.cs[6] = s1.z; kg.cs[7] = s1.w;
    const double4 bx = MODE != 2 ? a.box[id] : make_double4(0, 0, 0, 0);
    const double4 held = MODE == 2 ? a.com[id] : make_double4(0, 0, 0, 0);
    return internal_moments<MODE>(a, id, kg, bx, held);
}

constexpr int MOMENT_TOP_TB = 512, MOMENT_TOP_MAX = 2048;
constexpr int MOMENT_DEPTHS = 32;            // node depths are at most GF_KEY_LEVELS
static_assert(GF_KEY_LEVELS < MOMENT_DEPTHS, "depth bins");

template <int MODE>
__global__ __launch_bounds__(MOMENT_TOP_TB) void k_moments_top(MomentArgs a,
                                                               const int* __restrict__ top,
                                                               int* __restrict__ ntop,
                                                               int top_max) {
    constexpr int K = MOMENT_TOP_MAX / MOMENT_TOP_TB;
    __shared__ int s_id[MOMENT_TOP_MAX];         // listed nodes by depth
    __shared__ int4 s_cs[MOMENT_TOP_MAX][2];     // their children, octant order
    __shared__ double4 s_bx[MOMENT_TOP_MAX];     // box (modes 0, 1) or held centre of mass
    __shared__ int s_cnt[MOMENT_DEPTHS], s_off[MOMENT_DEPTHS], s_pos[MOMENT_DEPTHS];
    __shared__ int s_dmax;
    const int nt = *ntop;
    const int t = threadIdx.x;
    if (t < MOMENT_DEPTHS) s_cnt[t] = 0;
    if (t == 0) s_dmax = 0;
    __syncthreads();
    if (nt == 0) {
        // nothing above the subtrees
    } else if (nt <= top_max) {
        // counting sort by depth; every load of a round goes out together
        int id[K], dep[K];
#pragma unroll
        for (int k = 0; k < K; ++k) id[k] = top[min(t + MOMENT_TOP_TB * k, nt - 1)];
#pragma unroll
        for (int k = 0; k < K; ++k) dep[k] = a.depth[id[k]];
        int md = 0;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (t + MOMENT_TOP_TB * k < nt) {
                atomicAdd(&s_cnt[dep[k]], 1);
                md = max(md, dep[k]);
            }
        atomicMax(&s_dmax, md);
        __syncthreads();
        if (t == 0) {
            int o = 0;
            for (int d = 0; d < MOMENT_DEPTHS; ++d) {
                s_off[d] = s_pos[d] = o;
                o += s_cnt[d];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (t + MOMENT_TOP_TB * k < nt) s_id[atomicAdd(&s_pos[dep[k]], 1)] = id[k];
        __syncthreads();
        int4 c0[K], c1[K];
        double4 bx[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int nid = s_id[min(t + MOMENT_TOP_TB * k, nt - 1)];
            const int4* sl = reinterpret_cast<const int4*>(a.child8 + 8 * size_t(nid));
            c0[k] = sl[0];
            c1[k] = sl[1];
            bx[k] = MODE != 2 ? a.box[nid] : a.com[nid];
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int e = t + MOMENT_TOP_TB * k;
            if (e < nt) {
                s_cs[e][0] = c0[k];
                s_cs[e][1] = c1[k];
                s_bx[e] = bx[k];
            }
        }
        __syncthreads();
        for (int d = s_dmax; d >= 0; --d) {
            const int e1 = s_off[d] + s_cnt[d];
            for (int e = s_off[d] + t; e < e1; e += MOMENT_TOP_TB) {
                const int4 q0 = s_cs[e][0], q1 = s_cs[e][1];
                const KidsGlobal kg{a, s_id[e], {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w}};
                const double4 v = s_bx[e];
                (void)internal_moments<MODE>(a, s_id[e], kg, v, v);
            }
            __syncthreads();
        }
    } else {
        int md = 0;
        for (int k = t; k < nt; k += MOMENT_TOP_TB) md = max(md, int(a.depth[top[k]]));
        atomicMax(&s_dmax, md);
        __syncthreads();
        for (int d = s_dmax; d >= 0; --d) {
            for (int k = t; k < nt; k += MOMENT_TOP_TB) {
                const int id = top[k];
                if (a.depth[id] == d) (void)top_node<MODE>(a, id);
            }
            __syncthreads();
        }
    }
    if (t == 0) *ntop = 0;
}

// Groups (nbody.cpp:234-243): the outermost nodes holding <= 64 particles, or
// leaves, i.e. those whose parent holds more than 64.
// in one single-pass launch over the node ids: the group nodes in DFS
// order to groups[], their number to counters[1]
struct GroupOp {
    const int* count;
    const int* parent;
    const int8_t* leaf;
    int* groups;
    int* counters;
    __device__ void values(int i, int n, int (&v)[4]) const {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int id = i + q;
            bool g = false;
            if (id < n) {
                const int p = parent[id];
                g = (count[id] <= GF_GROUP_SIZE || leaf[id]) &&
                    (p < 0 || count[p] > GF_GROUP_SIZE);
            }
            v[q] = g;
        }
    }
    __device__ void put(int i, int n, const int (&v)[4], const int (&o)[4]) const {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i + q < n && v[q]) groups[o[q]] = i + q;
    }
    __device__ void total(int, int s) const { counters[1] = s; }
};

template <typename T>
int dalloc(T*& p, size_t count) {
    if (p) GF_HIP(hipFree(p));
    p = nullptr;
    GF_HIP(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (count ? count : 1)));
    return 0;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int node_grid(int cap_nodes) { return max(1, min(2048, gf_div_up(cap_nodes, TB))); }

// One single-pass launch of op over at most n_max items (len: the device
// count that bounds them, or null) with the look-back state in temp.
template <class Op>
int lookback_launch(const Op& op, int n_max, const int* len, int* temp, hipStream_t s) {
    unsigned int* timeouts = gf_prim_timeout_counter();
    if (!timeouts) {
        gf_set_error("tree build: no device address for the look-back counter");
        return -1;
    }
    hipLaunchKernelGGL(gf_k_lookback<Op>, dim3(gf_lb_blocks(n_max)), dim3(GF_LB_THREADS), 0, s,
                       op, n_max, len, reinterpret_cast<GfLbCtl*>(temp),
                       reinterpret_cast<unsigned long long*>(temp + GF_LB_CTL_INTS), timeouts);
    GF_LAUNCH_CHECK();
    return 0;
}

// Work items of several groups.  Groups (the largest nodes of at most 64
// particles) average about 18 particles on the disk and cloud workloads, and
// every group pays for a tree walk of its own; an item of several groups
// shares it.  An item is a run of sibling
// groups (children of one node, consecutive in octant order with only empty
// octants between them) taken in order while the item holds at most `cap`
// members; any other group is an item of its own.  The groups of an item
// are consecutive group positions and its members consecutive tree
// positions.  One thread per group; the first group of an item appends it
// (the order of the items changes no result: work on an item depends only
// on the item).
__global__ __launch_bounds__(TB) void k_merge_groups(const int* __restrict__ counters,
                                                  const int* __restrict__ groups,
                                                  const int* __restrict__ parent,
                                                  const int* __restrict__ child8,
                                                  const int* __restrict__ start,
                                                  const int* __restrict__ count, int cap,
                                                  int* __restrict__ nitems,
                                                  int4* __restrict__ items) {
    const int ng = counters[1];
    for (int gp = blockIdx.x * TB + threadIdx.x; gp < ng; gp += gridDim.x * TB) {
        const int G = groups[gp];
        const int np = count[G];
        const int P = parent[G];
        bool head = true;
        int ig = 1, im = np;
        if (np <= cap && P >= 0) {
            // the parent holds more than a group (else it would be the
            // group): its children of at most cap particles are groups, and
            // a larger child ends a run (its subtree's groups lie between)
            const int4* cs = reinterpret_cast<const int4*>(child8 + 8 * size_t(P));
            const int4 lo = cs[0], hi = cs[1];
            const int ch[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            int cnt[8];
#pragma unroll
            for (int o = 0; o < 8; ++o) cnt[o] = ch[o] >= 0 ? count[ch[o]] : 0;
            int cur = 0, first = -1, groups_in = 0;   // the open item
            bool open = false;
            for (int o = 0; o < 8; ++o) {
                if (ch[o] < 0) continue;              // empty octant: no particles between
                const bool fits = cnt[o] <= cap;
                if (open && (!fits || cur + cnt[o] > cap)) {
                    if (first == G) break;            // G's item ends before child o
                    open = false;
                }
                if (!fits) continue;
                if (!open) {
                    open = true;
                    cur = 0;
                    first = ch[o];
                    groups_in = 0;
                }
                cur += cnt[o];
                ++groups_in;
                if (ch[o] == G) head = first == G;
            }
            // G heads its item: cur and groups_in are that item's
            if (head) {
                ig = groups_in;
                im = cur;
            }
        }
        if (head) items[atomicAdd(nitems, 1)] = make_int4(gp, ig, start[G], im);
    }
}

// Subset trees: a stable compaction, in one single-pass launch, of the
// positions of the source tree's sorted order that hold one of the first m
// particles (their key and input index); the root box is the source tree's.
struct SubsetOp {
    const uint64_t* skey;
    const uint32_t* sidx;
    int m;
    const double* root_src;
    uint64_t* key;
    uint32_t* idx;
    double* root;
    __device__ void values(int i, int n, int (&v)[4]) const {
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = i + q < n && int(sidx[i + q]) < m;
    }
    __device__ void put(int i, int n, const int (&v)[4], const int (&o)[4]) const {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (i + q < n && v[q] && o[q] < m) {
                key[o[q]] = skey[i + q];
                idx[o[q]] = sidx[i + q];
            }
        }
    }
    __device__ void total(int, int) const {
        for (int c = 0; c < 4; ++c) root[c] = root_src[c];
    }
};

}  // namespace

int gf_tree_node_bound(int n) {
    return int(std::min<long long>(0x7fffffffLL,
                                   (long long)n + (long long)(GF_KEY_LEVELS + 1) * (n / (GF_LEAF_SIZE + 1)) + 2));
}

int gf_tree_reserve(GfTree& T, int n) {
    if (!T.counters) {
        if (dalloc(T.counters, 4) || dalloc(T.root, 4) || dalloc(T.bbox_part, 6 * 1024) ||
            dalloc(T.dmeta, 96))
            return -1;
    }
    if (n > T.cap_n) {
        const int cn = std::max(n, 1024);
        if (dalloc(T.key, cn) || dalloc(T.key_alt, cn) || dalloc(T.idx, cn) || dalloc(T.idx_alt, cn) ||
            dalloc(T.lcp, cn) || dalloc(T.d0, cn) || dalloc(T.chain, size_t(cn) + 1) ||
            dalloc(T.tx, cn) || dalloc(T.ty, cn) || dalloc(T.tz, cn) || dalloc(T.tm, cn) ||
            dalloc(T.tdisp, cn) || dalloc(T.xb, 3 * size_t(cn)) ||
            dalloc(T.scan_tmp, gf_scan_temp_ints(std::max(cn, gf_tree_node_bound(cn)) + 1)) ||
            gf_scan_temp_clear(T.scan_tmp, std::max(cn, gf_tree_node_bound(cn)) + 1) ||
            dalloc(T.sort_tmp, gf_sort_temp_ints(cn)))
            return -1;
        T.cap_n = cn;
        T.built = 0;
        if (T.skey) {
            dfree(T.skey);
            dfree(T.sidx);
        }
    }
    if (T.keep_sorted && !T.skey && (dalloc(T.skey, T.cap_n) || dalloc(T.sidx, T.cap_n)))
        return -1;
    const int nb = gf_tree_node_bound(T.cap_n);
    if (nb > T.cap_nodes) {
        if (dalloc(T.start, nb) || dalloc(T.count, nb) || dalloc(T.next, nb) ||
            dalloc(T.parent, nb) || dalloc(T.leaf, nb) || dalloc(T.depth, nb) ||
            dalloc(T.child8, 8 * size_t(nb)) || dalloc(T.dlist, nb) || dalloc(T.com, nb) ||
            dalloc(T.open_r, nb) || dalloc(T.box, nb) || dalloc(T.ndisp, nb) ||
            dalloc(T.quad, 6 * size_t(nb)) || dalloc(T.groups, nb))
            return -1;
        T.cap_nodes = nb;
        T.built = 0;
    }
    return 0;
}

void gf_tree_free(GfTree& T) {
    dfree(T.key); dfree(T.key_alt); dfree(T.idx); dfree(T.idx_alt); dfree(T.lcp); dfree(T.d0);
    dfree(T.chain); dfree(T.tx); dfree(T.ty); dfree(T.tz); dfree(T.tm); dfree(T.tdisp);
    dfree(T.xb); dfree(T.start); dfree(T.count); dfree(T.next); dfree(T.parent); dfree(T.leaf);
    dfree(T.depth); dfree(T.child8); dfree(T.dlist); dfree(T.dmeta); dfree(T.com); dfree(T.open_r);
    dfree(T.box); dfree(T.ndisp); dfree(T.quad); dfree(T.groups); dfree(T.counters);
    dfree(T.root); dfree(T.bbox_part); dfree(T.scan_tmp); dfree(T.sort_tmp);
    dfree(T.skey); dfree(T.sidx); dfree(T.sub_tmp);
    const int keep = T.keep_sorted;
    T = GfTree();
    T.keep_sorted = keep;
}

static MomentArgs moment_args(const GfTree& T) {
    MomentArgs a;
    a.start = T.start; a.count = T.count; a.next = T.next; a.parent = T.parent;
    a.leaf = T.leaf; a.depth = T.depth; a.child8 = T.child8; a.box = T.box;
    a.tx = T.tx; a.ty = T.ty; a.tz = T.tz; a.tm = T.tm; a.tdisp = T.tdisp;
    a.com = T.com; a.open_r = T.open_r; a.ndisp = T.ndisp; a.quad = T.quad;
    a.theta_d = double(T.theta);
    return a;
}

// The bottom-up pass over the held topology in two launches (subtrees, then
// the nodes above them); mode as leaf_moments.
static int moments_pass(const GfTree& T, int mode, hipStream_t s) {
    const MomentArgs a = moment_args(T);
    const int ng = node_grid(T.cap_nodes);
    int* ntop = T.dmeta;
    const int tl = gf_test_limits().moment_top;
    const int top_max = tl > 0 ? std::min(tl, MOMENT_TOP_MAX) : MOMENT_TOP_MAX;
    if (mode == 0) {
        hipLaunchKernelGGL(k_moments_subtrees<0>, dim3(ng), dim3(TB), 0, s, a, T.counters, T.dlist, ntop);
        hipLaunchKernelGGL(k_moments_top<0>, dim3(1), dim3(MOMENT_TOP_TB), 0, s, a, T.dlist, ntop,
                           top_max);
    } else if (mode == 1) {
        hipLaunchKernelGGL(k_moments_subtrees<1>, dim3(ng), dim3(TB), 0, s, a, T.counters, T.dlist, ntop);
        hipLaunchKernelGGL(k_moments_top<1>, dim3(1), dim3(MOMENT_TOP_TB), 0, s, a, T.dlist, ntop,
                           top_max);
    } else {
        hipLaunchKernelGGL(k_moments_subtrees<2>, dim3(ng), dim3(TB), 0, s, a, T.counters, T.dlist, ntop);
        hipLaunchKernelGGL(k_moments_top<2>, dim3(1), dim3(MOMENT_TOP_TB), 0, s, a, T.dlist, ntop,
                           top_max);
    }
    GF_LAUNCH_CHECK();
    return 0;
}

// Centres of mass (build and refit); the quadrupoles ride along when the
// last walk of this tree wanted them.
static int com_pass(GfTree& T, hipStream_t s) {
    const bool fuse = T.want_quad != 0;
    if (moments_pass(T, fuse ? 1 : 0, s)) return -1;
    T.quad_ok = fuse ? 1 : 0;
    return 0;
}

int gf_tree_merge_groups(const GfTree& T, int cap, int* nitems, int4* items, hipStream_t s) {
    if (T.n == 0) return 0;
    if (cap < 1 || cap > 64) {
        gf_set_error("merged groups: at most %d members, must be in [1, 64]", cap);
        return -1;
    }
    hipLaunchKernelGGL(k_merge_groups, dim3(std::min(2048, gf_div_up(T.cap_nodes, TB))), dim3(TB),
                       0, s, T.counters, T.groups, T.parent, T.child8, T.start, T.count, cap,
                       nitems, items);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_tree_quadrupoles(const GfTree& T, hipStream_t s) {
    if (T.quad_ok || T.n == 0) return 0;
    if (moments_pass(T, 2, s)) return -1;
    T.quad_ok = 1;
    return 0;
}

static int build_sorted(GfTree& T, int n, const double* pos, const double* mass, hipStream_t s,
                        bool with_com);

int gf_tree_build(GfTree& T, int n, const double* pos, const double* mass, float theta,
                  hipStream_t s, bool with_com, hipEvent_t sorted_event) {
    if (gf_tree_reserve(T, n)) return -1;
    T.n = n;
    T.theta = theta;
    T.built = 0;
    T.quad_ok = 0;
    if (n == 0) return 0;
    const int g = gf_div_up(n, TB);
    const int nbb = std::min(1024, g);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nbb), dim3(TB), 0, s, pos, n, T.bbox_part, T.idx);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(TB), 0, s, T.bbox_part, nbb, T.root);
    hipLaunchKernelGGL(k_keys, dim3(g), dim3(TB), 0, s, pos, n, T.root, T.key);
    GF_LAUNCH_CHECK();
    int in_alt = 0;
    if (gf_radix_sort_pairs(T.key, T.idx, T.key_alt, T.idx_alt, n, 0, 3 * GF_KEY_LEVELS,
                            T.sort_tmp, s, &in_alt))
        return -1;
    if (in_alt) {
        std::swap(T.key, T.key_alt);
        std::swap(T.idx, T.idx_alt);
    }
    if (T.keep_sorted) {
        GF_HIP(hipMemcpyAsync(T.skey, T.key, sizeof(uint64_t) * size_t(n), hipMemcpyDeviceToDevice, s));
        GF_HIP(hipMemcpyAsync(T.sidx, T.idx, sizeof(uint32_t) * size_t(n), hipMemcpyDeviceToDevice, s));
        if (sorted_event) GF_HIP(hipEventRecord(sorted_event, s));
    }
    return build_sorted(T, n, pos, mass, s, with_com);
}

// The build from the Morton-sorted (key, input index) on: nodes, leaf order,
// gathered particles, centres of mass (with_com) and groups.
static int build_sorted(GfTree& T, int n, const double* pos, const double* mass, hipStream_t s,
                        bool with_com) {
    const int g = gf_div_up(n, TB);
    hipLaunchKernelGGL(k_lcp, dim3(g), dim3(TB), 0, s, T.key, n, T.lcp);
    GF_LAUNCH_CHECK();
    if (lookback_launch(ChainOp{T.lcp, T.chain, T.d0, T.counters}, n, nullptr, T.scan_tmp, s))
        return -1;
    const int ng = node_grid(T.cap_nodes);
    hipLaunchKernelGGL(k_node_init, dim3(ng), dim3(TB), 0, s, T.counters, T.child8, T.dmeta);
    hipLaunchKernelGGL(k_emit, dim3(g), dim3(TB), 0, s, T.key, T.chain, T.d0, n, T.root, T.start,
                       T.count, T.next, T.parent, T.leaf, T.depth, T.box, T.child8, T.counters);
    hipLaunchKernelGGL(k_leaf_sort, dim3(ng), dim3(TB), 0, s, T.counters, T.start, T.count,
                       T.leaf, T.idx, T.key);
    // the positions at the build and the displacements since then serve the
    // centres of mass and refits only
    if (with_com)
        GF_HIP(hipMemcpyAsync(T.xb, pos, sizeof(double) * 3 * size_t(n), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_gather, dim3(g), dim3(TB), 0, s, pos, mass, T.idx, n, nullptr, T.tx,
                       T.ty, T.tz, T.tm, with_com ? T.tdisp : nullptr);
    GF_LAUNCH_CHECK();
    if (with_com && com_pass(T, s)) return -1;
    // groups: the node count lives on the device (counters[0]), so the launch
    // covers the node bound of this n and only the node count's tiles take part
    const int len = gf_tree_node_bound(n);
    if (lookback_launch(GroupOp{T.count, T.parent, T.leaf, T.groups, T.counters}, len,
                        T.counters, T.scan_tmp, s))
        return -1;
    T.built = 1;
    T.refit_ok = with_com;
    ++T.version;
    return 0;
}

int gf_tree_build_subset(GfTree& S, const GfTree& T, int m, const double* pos, const double* mass,
                         hipStream_t s) {
    if (m < 0 || m > T.n) {
        gf_set_error("subset tree: m = %d must be in [0, %d]", m, T.n);
        return -1;
    }
    if (!T.keep_sorted || !T.skey || !T.built) {
        gf_set_error("subset tree: the source tree keeps no sorted keys");
        return -1;
    }
    if (gf_tree_reserve(S, m)) return -1;
    S.n = m;
    S.theta = T.theta;
    S.built = 0;
    S.quad_ok = 0;
    if (m == 0) return 0;
    const int n = T.n;
    if (n + 1 > S.cap_sub) {
        if (dalloc(S.sub_tmp, gf_scan_temp_ints(n + 1)) || gf_scan_temp_clear(S.sub_tmp, n + 1))
            return -1;
        S.cap_sub = n + 1;
    }
    if (lookback_launch(SubsetOp{T.skey, T.sidx, m, T.root, S.key, S.idx, S.root}, n, nullptr,
                        S.sub_tmp, s))
        return -1;
    return build_sorted(S, m, pos, mass, s, false);
}

int gf_tree_refit(GfTree& T, const double* pos, const double* mass, hipStream_t s) {
    const int n = T.n;
    if (!T.built) {
        gf_set_error("refit without a built tree");
        return -1;
    }
    if (!T.refit_ok) {
        gf_set_error("refit of a tree built without centres of mass");
        return -1;
    }
    if (n == 0) return 0;
    const int g = gf_div_up(n, TB);
    const int ng = node_grid(T.cap_nodes);
    hipLaunchKernelGGL(k_gather, dim3(g), dim3(TB), 0, s, pos, mass, T.idx, n, T.xb, T.tx, T.ty,
                       T.tz, T.tm, T.tdisp);
    GF_LAUNCH_CHECK();
    (void)ng;
    ++T.version;
    return com_pass(T, s);
}
