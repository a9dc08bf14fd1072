// SPH gas physics on the device: exact kNN, smoothing lengths, density and
// pressure/viscosity forces (galaxy_formation/sph.py, simulation.py:74-95).
//
// Neighbour solve.  The reference takes the k nearest gas particles of each
// gas particle (itself included) from a k-d tree (sph.py:165-166) and sets
// h = max(d_k / 2, eps / 2) (simulation.py:84).  Here the gas particles get an
// octree of their own (the gravity builder), and the solve is a radius
// iteration: each particle owns a search radius r_i; a pass gathers the gas
// particles inside r_i and, when at least kq are inside, selects the exact kq
// nearest; otherwise r_i grows and the particle takes another pass.  The
// fast pass (k_knn_fast) serves one group of <= 64 gas particles per block of
// four waves; groups it cannot finish go to a heap pass (k_knn_pass: one wave
// per group, a max-heap of the kq best (d^2, input index) per lane in LDS).
// A solve asked for sorted rows sorts each row at the end (k_sort_rows).  Distances are FP64 with the
// operations of the oracle (no contraction), so neighbour sets and d_k are
// bit-identical to the brute-force oracle whatever the first radius.
//
// Density and forces.  One lane per gas particle in Morton order, the
// neighbour table stored column-major ([k][n]) so a wave reads 64 adjacent
// particles' j-th neighbours together; each neighbour's data is one packed
// 128-byte record, mostly an L2 hit.
#include "gf_prim.h"
#include "gf_sph.h"

namespace {

constexpr int TB = 256;
constexpr int CH = 128;                                 // candidate chunk (LDS)
constexpr int KSCAP = 1024;                             // LDS stack entries
constexpr int KGSCAP = 512 * (GF_KEY_LEVELS + 2);       // global stack entries per block
constexpr int KNN_BLOCKS = 4096;
constexpr int KNN_GST_BLOCKS = 128;
constexpr int GF_ANC = 32;                             // ancestors held per group (depth <= 21)

// A gas-tree node packed for the neighbour walks (64 bytes): the box, then
// the 8 children in octant order (-1 = empty), or, for a node of at most
// KNN_RANGE particles (contiguous in tree order) or a leaf, its particle
// range (KNODE_RANGE, first, count, depth-capped leaf of > GF_LEAF_SIZE).
constexpr int KNODE_RANGE = -2;
constexpr int KNN_RANGE = 64;
struct KNode {
    double4 box;
    int4 lo, hi;
};

struct KnnArgs {
    const KNode* kn;
    const double4* p4;      // (x, y, z, -) per tree position
    const int* start;
    const int* count;
    const int8_t* leaf;
    const int* child8;
    const double4* box;
    const int* groups;
    const int* counters;
    const double* tx;
    const double* ty;
    const double* tz;
    const uint32_t* idx;
    const int* inv;
    const int* parent_;
    const int* leaf_of;
    const int* anc;         // per group position: the group node and its ancestors
    double* r;
    double* r_lo;   // largest radius known to hold fewer than kq particles (0 = none)
    double* r_hi;   // smallest radius known to overflow the member list (inf = none)
    double* dk;
    int8_t* done;
    int* passes;
    int* nb;
    int n, kq;
    int tree_pos_out;
    int* state;     // [0] unresolved, [1] cursor of the fast pass, [2] cursor of the
                    // heap pass, [3] queue 1 length, [4] errors, [5] cursor of the
                    // global-stack heap pass, [6] queue 2 length, [7] retry list
                    // length; [8] groups of this pass's retry list; [9], [10]
                    // queue 1 and 2 lengths summed over passes; [11] passes that
                    // had work; [12] errors summed over passes; [13] internal
                    // bounds errors of the fast pass (bit 0 candidate range,
                    // bit 1 member-range stack); [14] merged work items
    int* ovf;       // queue 1: groups the fast pass hands to the heap pass
    int* ovf2;      // queue 2: groups the heap pass hands to its global-stack variant
    int* gstack;
    const int* retry_in;    // groups of this pass (null = all; count in state[8])
    int* retry_out;         // groups that need another pass (count in state[7])
    int* gmark;             // pass tag of the last retry request per group
    int pass_tag;
    int list_cap;           // candidate-list capacity (tests lower it)
    const int4* items;      // first pass only (null = one group per item): merged work items
                            // (first group position, groups, first particle, members); their
                            // count is state[14]
};

// Nodes to pop from a LIFO of capacity cap holding sn entries: a batch of 64
// while 64 nodes' children fit with room to spare, else one (depth first).
// Depth first from any state needs at most 7 more entries per tree level, so
// a reserve of 7 * 64 + 8 * (GF_KEY_LEVELS + 1) keeps every push in bounds.
__device__ __forceinline__ int stack_take(int sn, int cap) {
    return (cap - sn >= 7 * 64 + 8 * (GF_KEY_LEVELS + 1)) ? min(64, sn) : 1;
}

// Group position gp needs another pass (once per pass, whoever asks).
__device__ __forceinline__ void request_retry(const KnnArgs& a, int gp) {
    if (atomicExch(&a.gmark[gp], a.pass_tag) != a.pass_tag)
        a.retry_out[atomicAdd(&a.state[7], 1)] = gp;
}

__device__ __forceinline__ bool key_less(double da, int ia, double db, int ib) {
    return da < db || (da == db && ia < ib);
}

template <int KQM, bool GST>
__global__ __launch_bounds__(64) void k_knn_pass(KnnArgs a) {
#pragma clang fp contract(off)
    __shared__ double hd[KQM * 64];
    __shared__ int hi[KQM * 64];
    __shared__ double cx[CH], cy[CH], cz[CH];
    __shared__ int ci[CH];
    __shared__ int s_stack[GST ? 1 : KSCAP];
    const int lane = __lane_id();
    int* stack = GST ? a.gstack + size_t(blockIdx.x) * KGSCAP : s_stack;
    const int scap = GST ? KGSCAP : KSCAP;
    const int* list = GST ? a.ovf2 : a.ovf;
    const int ngroups = GST ? a.state[6] : a.state[3];
    if (ngroups == 0) return;          // nothing queued (the common case)
    int* cursor = a.state + (GST ? 5 : 2);
    const int kq = a.kq;

    while (true) {
        int gi = 0;
        if (lane == 0) gi = atomicAdd(cursor, 1);
        gi = __shfl(gi, 0, 64);
        if (gi >= ngroups) break;
        const int gp = list[gi];
        const int G = a.groups[gp];
        const int b0 = a.start[G], np = a.count[G];
        for (int c0 = 0; c0 < np; c0 += 64) {
            const int t = b0 + c0 + lane;
            const bool in = c0 + lane < np;
            const bool mine = in && !a.done[t];
            if (__ballot(mine) == 0) continue;
            double xi = 0, yi = 0, zi = 0, ri = 0;
            if (in) {
                xi = a.tx[t]; yi = a.ty[t]; zi = a.tz[t];
            }
            if (mine) ri = a.r[t];
            const double r2i = ri * ri;
            // sphere (b, R) around every unresolved member's search sphere
            double x0 = mine ? xi : 1e300, y0 = mine ? yi : 1e300, z0 = mine ? zi : 1e300;
            double x1 = mine ? xi : -1e300, y1 = mine ? yi : -1e300, z1 = mine ? zi : -1e300;
            x0 = gf_wave_min_d(x0); y0 = gf_wave_min_d(y0); z0 = gf_wave_min_d(z0);
            x1 = gf_wave_max_d(x1); y1 = gf_wave_max_d(y1); z1 = gf_wave_max_d(z1);
            const double bx = 0.5 * (x0 + x1), by = 0.5 * (y0 + y1), bz = 0.5 * (z0 + z1);
            double reach = 0;
            if (mine) {
                const double dx = xi - bx, dy = yi - by, dz = zi - bz;
                reach = sqrt(dx * dx + dy * dy + dz * dz) + ri;
            }
            const double R = gf_wave_max_d(reach) * (1.0 + 1e-12) + 1e-300;
            const double R2 = R * R;

            int hsize = 0, cnt = 0;
            int sn = 1, cn = 0;
            bool overflow = false;
            if (lane == 0) stack[0] = 0;
            __syncthreads();

            // candidates in the chunk enter the heaps of the lanes whose
            // search sphere holds them
            auto flush = [&]() {
                __syncthreads();
                if (mine) {
                    for (int c = 0; c < cn; ++c) {
                        const double dx = xi - cx[c], dy = yi - cy[c], dz = zi - cz[c];
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 > r2i) continue;
                        ++cnt;
                        const int id = ci[c];
                        if (hsize < kq) {
                            int p = hsize++;
                            while (p > 0) {
                                const int q = (p - 1) >> 1;
                                const double qd = hd[q * 64 + lane];
                                const int qi = hi[q * 64 + lane];
                                if (!key_less(qd, qi, d2, id)) break;
                                hd[p * 64 + lane] = qd;
                                hi[p * 64 + lane] = qi;
                                p = q;
                            }
                            hd[p * 64 + lane] = d2;
                            hi[p * 64 + lane] = id;
                        } else if (key_less(d2, id, hd[lane], hi[lane])) {
                            int p = 0;
                            while (true) {
                                int q = 2 * p + 1;
                                if (q >= kq) break;
                                double qd = hd[q * 64 + lane];
                                int qi = hi[q * 64 + lane];
                                if (q + 1 < kq) {
                                    const double rd = hd[(q + 1) * 64 + lane];
                                    const int rr = hi[(q + 1) * 64 + lane];
                                    if (key_less(qd, qi, rd, rr)) {
                                        ++q;
                                        qd = rd;
                                        qi = rr;
                                    }
                                }
                                if (!key_less(d2, id, qd, qi)) break;
                                hd[p * 64 + lane] = qd;
                                hi[p * 64 + lane] = qi;
                                p = q;
                            }
                            hd[p * 64 + lane] = d2;
                            hi[p * 64 + lane] = id;
                        }
                    }
                }
                cn = 0;
                __syncthreads();
            };

            while (sn > 0) {
                const int take = stack_take(sn, scap);
                sn -= take;
                const int node = lane < take ? stack[sn + lane] : -1;
                __syncthreads();
                int push = 0, nent = 0, first = 0;
                int4 c_lo = make_int4(-1, -1, -1, -1), c_hi = c_lo;
                if (node >= 0) {
                    const double4 bb = a.box[node];
                    const double ex = fmax(0.0, fabs(bx - bb.x) - bb.w);
                    const double ey = fmax(0.0, fabs(by - bb.y) - bb.w);
                    const double ez = fmax(0.0, fabs(bz - bb.z) - bb.w);
                    if (ex * ex + ey * ey + ez * ez <= R2) {
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
                }
                const int pinc = gf_wave_incl_scan(push);
                const int ptot = __builtin_amdgcn_readlane(pinc, 63);
                if (sn + ptot > scap) {
                    overflow = true;
                    break;
                }
                {
                    int o = sn + pinc - push;
                    const int cc[8] = {c_lo.x, c_lo.y, c_lo.z, c_lo.w, c_hi.x, c_hi.y, c_hi.z, c_hi.w};
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                        if (cc[k] >= 0) stack[o++] = cc[k];
                }
                sn += ptot;
                int rem = nent, src = 0;
                while (true) {
                    const int einc = gf_wave_incl_scan(rem);
                    const int etot = __builtin_amdgcn_readlane(einc, 63);
                    if (etot == 0) break;
                    if (cn == CH) flush();
                    const int space = CH - cn;
                    const int my0 = einc - rem;
                    const int canw = max(0, min(rem, space - my0));
                    for (int k = 0; k < canw; ++k) {
                        const int slot = cn + my0 + k;
                        const int j = first + src + k;
                        cx[slot] = a.tx[j];
                        cy[slot] = a.ty[j];
                        cz[slot] = a.tz[j];
                        ci[slot] = int(a.idx[j]);
                    }
                    rem -= canw;
                    src += canw;
                    cn += min(etot, space);
                }
                __syncthreads();
            }
            if (overflow) {
                if (!GST) {
                    if (lane == 0) a.ovf2[atomicAdd(&a.state[6], 1)] = gp;
                } else if (lane == 0) {
                    atomicAdd(&a.state[4], 1);
                }
                __syncthreads();
                break;
            }
            if (cn > 0) flush();

            if (mine) {
                a.passes[t] += 1;
                if (cnt >= kq) {
                    a.dk[t] = sqrt(hd[lane]);            // the heap's top: the kq-th
                    for (int s = 0; s < kq; ++s) {
                        const int id = hi[s * 64 + lane];
                        a.nb[size_t(s) * a.n + t] = a.tree_pos_out ? a.inv[id] : id;
                    }
                    a.done[t] = 1;
                } else {
                    // plain iteration: grow the sphere towards kq particles
                    const double want = double(kq) / double(max(cnt, 1));
                    const double f = fmin(2.0, fmax(1.2, 1.1 * cbrt(want)));
                    a.r[t] = ri * f;
                    atomicAdd(&a.state[0], 1);
                }
            }
            if (__ballot(mine && cnt < kq) != 0 && lane == 0) request_retry(a, gp);
            __syncthreads();
        }
    }
}

// Between passes: this pass visits the previous pass's retry list (all
// groups on the first pass); the per-pass counters restart; totals carry on.
__global__ void k_knn_reset(int* state, int first) {
    if (threadIdx.x != 0) return;
    if (first) {
        state[9] = state[10] = state[11] = state[12] = state[13] = 0;
    } else {
        state[9] += state[3];
        state[10] += state[6];
        state[12] += state[4];
    }
    const int work = first ? 1 : state[7];
    state[8] = first ? 0 : state[7];
    state[11] += work > 0;
    for (int i = 0; i < 8; ++i) state[i] = 0;
}

// ----------------------------------------------------- fast neighbour pass
// One block of four waves per gas group.  Candidates (gas particles inside a
// search sphere) are gathered into LDS by a tree walk over packed nodes that
// starts at the smallest node whose box strictly contains the sphere.
// Per member the 64 lanes of a wave compute the exact FP64 squared distances
// to the list, count those within the member's radius, select the kq
// nearest (radix rounds on fixed-point keys, then an exact rank of the
// boundary bucket by (d^2, index)) and write the row unordered; d_k is the
// kq-th exact distance.  A member with fewer than kq inside its radius grows
// the radius; one whose sphere holds more candidates than the list shrinks
// it; both take another pass.
constexpr int FT = 256;
// Two shapes of the fast pass.  A solve that starts from the caller's radii
// (the previous d_k, the learned proposal) holds about k candidates per
// member: a 512-entry list, 4 blocks per CU, and the union of the member
// spheres gathered first.  The cell-density start overshoots: a 1024-entry
// list and the bounding sphere first (fewer shrink passes).
// five narrow blocks per CU (31 KB of LDS each, at most 102 VGPRs): the pass
// waits on memory most of its time, and the fifth block hides more of it
// (1.55 -> 1.39 ms alone on the 500k gas of the 1M disk)
#ifndef KNN_NARROW_BLOCKS_PER_CU
#define KNN_NARROW_BLOCKS_PER_CU 5
#endif
#ifndef KNN_NARROW_CMAX
#define KNN_NARROW_CMAX 512
#endif
struct KnnNarrow {
    static constexpr int CMAX = KNN_NARROW_CMAX, BLOCKS_PER_CU = KNN_NARROW_BLOCKS_PER_CU,
                         UNION_FIRST = 1;
};
struct KnnWide {
    static constexpr int CMAX = 1024, BLOCKS_PER_CU = 3, UNION_FIRST = 0;
};
constexpr int CMAX_MAX = 1024;               // largest candidate list (test limits)
constexpr int FSCAP = 1024;                  // stack of the gather walk
// persistent blocks of the fast pass by default: the narrow shape's
// residency (5 per CU on 256 CUs), and 1024 for the wide shape (3 per CU);
// more only queue behind the resident ones and hold back the gravity walk
// that runs beside the solve on another stream
#ifndef KNN_FAST_GRID
#define KNN_FAST_GRID (256 * KNN_NARROW_BLOCKS_PER_CU)
#endif
#ifndef KNN_WIDE_GRID
#define KNN_WIDE_GRID 1024
#endif
constexpr int FAST_BLOCKS = KNN_FAST_GRID;
// members of a merged work item at most (first pass from a caller's radii)
#ifndef KNN_MERGE_MEMBERS
#define KNN_MERGE_MEMBERS 40
#endif
constexpr int WIDE_BLOCKS = KNN_WIDE_GRID;
constexpr uint32_t KEY_NONE = 0xFFFFFFFFu;
constexpr double KEY_SCALE = 4294967294.0;   // keys d^2 / r^2 * KEY_SCALE < KEY_NONE
constexpr double GROUP_REACH = 3.0;          // gather once while R <= 3 max r_i

// Smallest ancestor of `node` (node included) whose box strictly contains
// the sphere; the root when none does.
__device__ int containing_node(const KnnArgs& a, int node, double bx, double by, double bz,
                               double R) {
#pragma clang fp contract(off)
    while (true) {
        const int p = a.parent_ != nullptr ? a.parent_[node] : -1;
        if (p < 0) return node;
        const double4 bb = a.box[node];
        const double lim = bb.w * (1.0 - 1e-12);
        if (fabs(bx - bb.x) + R < lim && fabs(by - bb.y) + R < lim && fabs(bz - bb.z) + R < lim)
            return node;
        node = p;
    }
}

// containing_node from the group node, from the table of its ancestors: lane
// k tests ancestor k (one round of loads instead of a chain), and the
// deepest one that contains the sphere, or the root, is the answer (the
// boxes grow towards the root, so the tests pass from some depth up).
__device__ int group_containing_node(const KnnArgs& a, int gp, double bx, double by, double bz,
                                     double R) {
#pragma clang fp contract(off)
    const int lane = __lane_id();
    const int node = lane < GF_ANC ? a.anc[size_t(gp) * GF_ANC + lane] : -1;
    bool ok = false;
    if (node >= 0) {
        const double4 bb = a.box[node];
        const double lim = bb.w * (1.0 - 1e-12);
        ok = fabs(bx - bb.x) + R < lim && fabs(by - bb.y) + R < lim && fabs(bz - bb.z) + R < lim;
    }
    const uint64_t okm = __ballot(ok), valid = __ballot(node >= 0);
    const int last = 63 - __clzll(valid);
    if (okm == 0 && last == GF_ANC - 1)      // a chain longer than the table: walk on
        return containing_node(a, __builtin_amdgcn_readlane(node, last), bx, by, bz, R);
    return __builtin_amdgcn_readlane(node, okm ? __ffsll((unsigned long long)okm) - 1 : last);
}

// One lane's visit of a packed node against the sphere (b, R2): a range
// node that meets it gives its particles (first, cnt), an internal node its
// children; capped marks a depth-capped leaf of more than GF_LEAF_SIZE.
struct NodeHit {
    int push, first, cnt, capped;
    int4 c_lo, c_hi;
};

__device__ __forceinline__ bool box_meets(const double4& box, double x, double y, double z,
                                          double r2) {
#pragma clang fp contract(off)
    const double ex = fmax(0.0, fabs(x - box.x) - box.w);
    const double ey = fmax(0.0, fabs(y - box.y) - box.w);
    const double ez = fmax(0.0, fabs(z - box.z) - box.w);
    return ex * ex + ey * ey + ez * ez <= r2;
}

// Does the point or box meet one of the nm spheres (x, y, z, r^2) of msph?
__device__ __forceinline__ bool box_meets_any(const double4& box, const double4* msph, int nm) {
    for (int q = 0; q < nm; ++q) {
        const double4 m = msph[q];
        if (box_meets(box, m.x, m.y, m.z, m.w)) return true;
    }
    return false;
}

__device__ __forceinline__ bool point_in_any(double x, double y, double z, const double4* msph,
                                             int nm) {
#pragma clang fp contract(off)
    for (int q = 0; q < nm; ++q) {
        const double4 m = msph[q];
        const double dx = x - m.x, dy = y - m.y, dz = z - m.z;
        if (dx * dx + dy * dy + dz * dz <= m.w) return true;
    }
    return false;
}

// One lane's visit of a packed node against the sphere (b, R2) and, when
// nm > 0, the union of the member spheres msph[0, nm).
__device__ __forceinline__ NodeHit visit_node(const KnnArgs& a, int node, double bx, double by,
                                              double bz, double R2, const double4* msph, int nm) {
#pragma clang fp contract(off)
    NodeHit h{0, 0, 0, 0, make_int4(-1, -1, -1, -1), make_int4(-1, -1, -1, -1)};
    if (node < 0) return h;
    const KNode kn = a.kn[node];
    if (box_meets(kn.box, bx, by, bz, R2) && (nm == 0 || box_meets_any(kn.box, msph, nm))) {
        if (kn.lo.x == KNODE_RANGE) {
            h.first = kn.lo.y;
            h.cnt = kn.lo.z;
            h.capped = kn.lo.w;
        } else {
            h.c_lo = kn.lo;
            h.c_hi = kn.hi;
            h.push = (kn.lo.x >= 0) + (kn.lo.y >= 0) + (kn.lo.z >= 0) + (kn.lo.w >= 0) +
                     (kn.hi.x >= 0) + (kn.hi.y >= 0) + (kn.hi.z >= 0) + (kn.hi.w >= 0);
        }
    }
    return h;
}

// Pushes the children of this batch's nodes at stack[sn ...] in lane order
// (the stack stays sorted by depth); returns how many.
__device__ __forceinline__ int push_children(const NodeHit& h, int* stack, int sn) {
    const int pinc = gf_wave_incl_scan(h.push);
    int o = sn + pinc - h.push;
    const int cc[8] = {h.c_lo.x, h.c_lo.y, h.c_lo.z, h.c_lo.w, h.c_hi.x, h.c_hi.y, h.c_hi.z, h.c_hi.w};
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (cc[k] >= 0) stack[o++] = cc[k];
    return __builtin_amdgcn_readlane(pinc, 63);
}

// The whole block gathers the particles within (b, R) of the subtree of
// `root`, and, when nm > 0, within one of the member spheres msph[0, nm)
// (x, y, z, r^2): wave 0 walks the tree and lists the ranges that meet them
// (lf: first particle, lincl: inclusive prefix of their counts); then every
// wave loads those ranges' particles and keeps the ones inside, in list
// order, so the candidate order is deterministic.  Called by all threads.
// Returns the count, -2 for a depth-capped leaf, -1 when the sphere holds
// more than cap particles (ctl[7] = an estimate of their number, from the
// share of entries inside) or its ranges more than CMAX / 4 CMAX (ctl[7] = 0).
template <int CMAX>
__device__ int gather_group(const KnnArgs& a, int root, double bx, double by, double bz,
                            double R2, const double4* msph, int nm, double* lx, double* ly,
                            double* lz, int* li, int cap, int* stack, int* lf, int* lincl,
                            int* ctl) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    const uint64_t lt = gf_lanemask_lt();
    if (w == 0) {
        int sn = 1, nl = 0, et = 0, res = 0;
        if (lane == 0) stack[0] = root;
        __builtin_amdgcn_wave_barrier();
        while (sn > 0) {
            const int take = stack_take(sn, FSCAP);
            sn -= take;
            const int node = lane < take ? stack[sn + lane] : -1;
            __builtin_amdgcn_wave_barrier();
            const NodeHit h = visit_node(a, node, bx, by, bz, R2, msph, nm);
            sn += push_children(h, stack, sn);
            if (__ballot(h.capped != 0) != 0) {
                res = -2;
                break;
            }
            const uint64_t lm = __ballot(h.cnt > 0);
            const int einc = gf_wave_incl_scan(h.cnt);
            const int ebatch = __builtin_amdgcn_readlane(einc, 63);
            const int nlb = __popcll(lm);
            // ranges touched at most CMAX, particles in them at most 4 CMAX
            if (nl + nlb > CMAX || et + ebatch > 4 * CMAX) {
                res = -1;
                break;
            }
            if (h.cnt > 0) {
                const int o = nl + __popcll(lm & lt);
                lf[o] = h.first;
                lincl[o] = et + einc;
            }
            nl += nlb;
            et += ebatch;
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) {
            ctl[4] = res;
            ctl[5] = nl;
            ctl[6] = et;
        }
    }
    __syncthreads();
    const int res = ctl[4], nl = ctl[5], et = ctl[6];
    if (res < 0) {
        if (threadIdx.x == 0) ctl[7] = 0;         // count not known
        __syncthreads();
        return res;
    }
    int nc = 0;
    for (int e0 = 0; e0 < et; e0 += FT) {
        const int e = e0 + tid;
        bool tk = false;
        double px = 0, py = 0, pz = 0;
        int pj = 0;
        if (e < et) {
            int lo = 0, hi = nl - 1;                 // first range with lincl > e
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (lincl[mid] > e) hi = mid; else lo = mid + 1;
            }
            pj = lf[lo] + (e - (lo > 0 ? lincl[lo - 1] : 0));
            if (pj >= 0 && pj < a.n) {
                const double4 q = a.p4[pj];
                px = q.x; py = q.y; pz = q.z;
                const double dx = px - bx, dy = py - by, dz = pz - bz;
                tk = dx * dx + dy * dy + dz * dz <= R2 &&
                     (nm == 0 || point_in_any(px, py, pz, msph, nm));
            } else {
                atomicOr(&a.state[13], 1);    // inconsistent range list: an error, no access
            }
        }
        const uint64_t m = __ballot(tk);
        if (lane == 0) ctl[8 + w] = __popcll(m);
        __syncthreads();
        int base = nc;
        for (int q = 0; q < w; ++q) base += ctl[8 + q];
        const int tot = ctl[8] + ctl[9] + ctl[10] + ctl[11];
        const int o = base + __popcll(m & lt);
        if (tk && o < cap) {
            lx[o] = px; ly[o] = py; lz[o] = pz;
            li[o] = pj;
        }
        nc += tot;
        __syncthreads();
        if (nc > cap) {
            // the list is full: estimate the sphere's count from the entries seen
            const long long seen = min(et, e0 + FT);
            nc = int(min(1ll << 30, (long long)nc * et / seen));
            break;
        }
    }
    if (threadIdx.x == 0) ctl[7] = nc;
    // every wave reads ctl[7] next (the chunk count of a split): the write
    // must land first
    __syncthreads();
    return nc > cap ? -1 : nc;
}

// Sphere (b, R) around the search spheres of the lanes with `mine`, and the
// largest of their radii (wave reductions; every lane gets the result).
struct Sphere {
    double bx, by, bz, R, rmax;
};

__device__ __forceinline__ Sphere wave_sphere(bool mine, double xi, double yi, double zi,
                                              double ri) {
#pragma clang fp contract(off)
    double x0 = mine ? xi : 1e300, y0 = mine ? yi : 1e300, z0 = mine ? zi : 1e300;
    double x1 = mine ? xi : -1e300, y1 = mine ? yi : -1e300, z1 = mine ? zi : -1e300;
    x0 = gf_wave_min_d(x0); y0 = gf_wave_min_d(y0); z0 = gf_wave_min_d(z0);
    x1 = gf_wave_max_d(x1); y1 = gf_wave_max_d(y1); z1 = gf_wave_max_d(z1);
    Sphere S;
    S.bx = 0.5 * (x0 + x1); S.by = 0.5 * (y0 + y1); S.bz = 0.5 * (z0 + z1);
    double reach = 0;
    if (mine) {
        const double dx = xi - S.bx, dy = yi - S.by, dz = zi - S.bz;
        reach = sqrt(dx * dx + dy * dy + dz * dz) + ri;
    }
    S.R = gf_wave_max_d(reach) * (1.0 + 1e-12) + 1e-300;
    S.rmax = gf_wave_max_d(mine ? ri : 0.0);
    return S;
}

// Exact kq nearest of member t (at (xi, yi, zi), radius ri) among the list
// (lx, ly, lz, li)[0, nc) (li: tree positions), by one wave.  Keys are
// d^2 / r^2 in 32-bit fixed point (monotone in the FP64 d^2).  Pass 1 forms
// the keys and the histogram of their top 8 bits; radix rounds of 8 bits
// narrow the bucket holding the kq-th key until it holds at most 64 keys;
// pass 2 writes the keys below that bucket to the row and gathers the
// bucket, whose entries are then ranked exactly by (d^2, input index), one
// per lane (input indices are read only when two of them tie in d^2).
// Returns 1 when resolved, 0 when the radius was too small (grown), -1 on
// more than 64 equal keys (left for the heap pass).
template <int CMAX>
__device__ int select_member(const KnnArgs& a, int t, double xi, double yi, double zi, double ri,
                             int npass, const double* lx, const double* ly, const double* lz,
                             const int* li, int nc, int* hist, double* tie_d, int* tie_i,
                             int* tie_p) {
#pragma clang fp contract(off)
    const int lane = __lane_id();
    const uint64_t lt = gf_lanemask_lt();
    const int kq = a.kq;
    const double r2i = ri * ri;
    const double ks = KEY_SCALE / r2i;
    const int ncl = (nc + 63) >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) hist[4 * lane + q] = 0;
    __builtin_amdgcn_wave_barrier();
    constexpr int CPL = CMAX / 64;
    uint32_t key[CPL];
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        uint32_t kk = KEY_NONE;
        if (c < ncl) {
            const int j = lane + 64 * c;
            if (j < nc) {
                const double dx = xi - lx[j], dy = yi - ly[j], dz = zi - lz[j];
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 <= r2i) {
                    kk = uint32_t(fmin(d2 * ks, KEY_SCALE));
                    atomicAdd(&hist[kk >> 24], 1);
                }
            }
        }
        key[c] = kk;
    }
    __builtin_amdgcn_wave_barrier();
    // radix rounds: `prefix` holds the bits above `shift` of the bucket of
    // the kq-th key, `need` its rank inside that bucket, `bcount` its size
    uint64_t prefix = 0;
    int need = kq, shift = 24, bcount = 0;
    while (true) {
        int h[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) h[q] = hist[4 * lane + q];
        const int sum = h[0] + h[1] + h[2] + h[3];
        const int incl = gf_wave_incl_scan(sum);
        if (shift == 24) {
            const int total = __builtin_amdgcn_readlane(incl, 63);
            if (total < kq) {
                if (lane == 0) {
                    a.passes[t] = npass + 1;
                    const double want = double(kq) / double(max(total, 1));
                    double rn = ri * fmin(2.0, fmax(1.2, 1.1 * cbrt(want)));
                    a.r_lo[t] = ri;
                    const double hi = a.r_hi[t];
                    if (rn >= hi) rn = sqrt(ri * hi);       // bisect inside the bracket
                    a.r[t] = rn;
                    atomicAdd(&a.state[0], 1);
                }
                return 0;
            }
        }
        const int excl = incl - sum;
        const bool here = excl < need && need <= incl;
        int bsel = 0, nneed = 0, bc = 0;
        if (here) {
            int run = excl;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (nneed == 0 && run + h[q] >= need) {
                    bsel = 4 * lane + q;
                    nneed = need - run;
                    bc = h[q];
                }
                run += h[q];
            }
        }
        const int src = __ffsll((unsigned long long)__ballot(here)) - 1;
        bsel = __builtin_amdgcn_readlane(bsel, src);
        need = __builtin_amdgcn_readlane(nneed, src);
        bcount = __builtin_amdgcn_readlane(bc, src);
        prefix |= uint64_t(bsel) << shift;
        __builtin_amdgcn_wave_barrier();
        if (bcount <= 64 || shift == 0) break;
        shift -= 8;
        reinterpret_cast<int4*>(hist)[lane] = make_int4(0, 0, 0, 0);
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
            const uint64_t kk = key[c];
            if (c < ncl && kk != KEY_NONE && (kk >> (shift + 8)) == (prefix >> (shift + 8)))
                atomicAdd(&hist[(kk >> shift) & 0xFF], 1);
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (bcount > 64) return -1;
    const uint64_t pfx = prefix >> shift;
    // pass 2: keys below the bucket go to the row, the bucket to the tie lists
    // (both compacted by ballots: slot by slot, lanes in order)
    const size_t n = size_t(a.n);
    int ob = 0, tb = 0;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
        if (c < ncl) {
            const uint64_t kd = uint64_t(key[c]) >> shift;
            const bool valid = key[c] != KEY_NONE;
            const bool isb = valid && kd < pfx;
            const bool ist = valid && kd == pfx;
            const uint64_t mb = __ballot(isb);
            const uint64_t mt = __ballot(ist);
            const int j = lane + 64 * c;
            if (isb) {
                const int pj = li[j];
                a.nb[size_t(ob + __popcll(mb & lt)) * n + t] = a.tree_pos_out ? pj : int(a.idx[pj]);
            }
            if (ist) {
                const int o = tb + __popcll(mt & lt);
                const double dx = xi - lx[j], dy = yi - ly[j], dz = zi - lz[j];
                tie_d[o] = dx * dx + dy * dy + dz * dz;
                tie_p[o] = li[j];
            }
            ob += __popcll(mb);
            tb += __popcll(mt);
        }
    }
    __builtin_amdgcn_wave_barrier();
    const bool bl = lane < bcount;
    double dm = 0;
    int pm = 0, rank = 0, same = 0;
    if (bl) {
        dm = tie_d[lane];
        pm = tie_p[lane];
        int q = 0;
        for (; q + 4 <= bcount; q += 4) {
            const double d0 = tie_d[q], d1 = tie_d[q + 1], d2 = tie_d[q + 2], d3 = tie_d[q + 3];
            rank += (d0 < dm) + (d1 < dm) + (d2 < dm) + (d3 < dm);
            same += (d0 == dm) + (d1 == dm) + (d2 == dm) + (d3 == dm);
        }
        for (; q < bcount; ++q) {
            const double dq = tie_d[q];
            rank += dq < dm;
            same += dq == dm;
        }
    }
    if (__ballot(same > 1) != 0) {
        // equal distances: the input index decides (sph.py:165-166 order)
        const int im = bl ? int(a.idx[pm]) : 0;
        if (bl) tie_i[lane] = im;
        __builtin_amdgcn_wave_barrier();
        if (bl && same > 1)
            for (int q = 0; q < bcount; ++q) rank += tie_d[q] == dm && tie_i[q] < im;
        __builtin_amdgcn_wave_barrier();
    }
    const bool take = bl && rank < need;
    const double dk2 = gf_wave_max_d(take ? dm : 0.0);
    if (take) {
        const int o = ob + __popcll(__ballot(take) & lt);
        a.nb[size_t(o) * n + t] = a.tree_pos_out ? pm : int(a.idx[pm]);
    }
    if (lane == 0) {
        a.dk[t] = sqrt(dk2);
        a.done[t] = 1;
        a.passes[t] = npass + 1;
    }
    return 1;
}

// One pass of the neighbour solve for the groups of this pass (all groups,
// or the previous pass's retry list).  A block of 4 waves takes one group.
// Its unresolved members are served by ranges of members (contiguous in
// Morton order, so spatially compact), starting with the whole group: the
// block gathers the sphere around the range's search spheres once, and each
// wave selects for every fourth member of the range.  A range whose sphere
// holds more than the list is halved; a single member whose own sphere does
// not fit shrinks its radius and takes another pass.  Groups it cannot
// finish (a depth-capped leaf, more than 64 equal keys, a closed radius
// bracket) go to the heap pass's queue.
template <typename Shape>
__global__ __launch_bounds__(FT, Shape::BLOCKS_PER_CU) void k_knn_fast(KnnArgs a) {
#pragma clang fp contract(off)
    constexpr int CMAX = Shape::CMAX;
    __shared__ double cx[CMAX], cy[CMAX], cz[CMAX];
    __shared__ int ci[CMAX];
    __shared__ int stack[FSCAP];
    __shared__ int lf[CMAX], lincl[CMAX];
    __shared__ __align__(16) int hist[4][256];
    __shared__ double tie_d[4][64];
    __shared__ int tie_i[4][64];
    __shared__ int tie_p[4][64];
    __shared__ int ctl[16];
    __shared__ unsigned long long rstack[16];
    __shared__ int rmode[16];                   // 0: bounding sphere, 1: union of member spheres
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    // the first pass serves merged work items when the solve built them; its
    // retries, the later passes and the heap pass's queue are groups
    const bool use_items = a.items != nullptr && a.retry_in == nullptr;
    const int nwork = a.retry_in ? a.state[8] : (use_items ? a.state[14] : a.counters[1]);
    const int cap = min(CMAX, a.list_cap);
    if (threadIdx.x < 16) ctl[threadIdx.x] = 0;
    __syncthreads();

    while (true) {
        if (threadIdx.x == 0) ctl[0] = atomicAdd(&a.state[1], 1);
        __syncthreads();
        const int gi = ctl[0];
        if (gi >= nwork) break;
        // the work's groups are positions [gp, gp + ngrp), its members the
        // tree positions [b0, b0 + np)
        int gp, ngrp, b0, np;
        if (use_items) {
            const int4 it = a.items[gi];
            gp = it.x; ngrp = it.y; b0 = it.z; np = it.w;
        } else {
            gp = a.retry_in ? a.retry_in[gi] : gi;
            ngrp = 1;
            const int G = a.groups[gp];
            b0 = a.start[G]; np = a.count[G];
        }
        if (np > 64) {
            // a depth-capped leaf of more than 64 (always an item of its own)
            if (threadIdx.x == 0)
                for (int g = 0; g < ngrp; ++g) a.ovf[atomicAdd(&a.state[3], 1)] = gp + g;
            __syncthreads();
            continue;
        }
        // members' data, lane m = member m (every wave holds the same)
        const int tm = b0 + lane;
        const bool inm = lane < np;
        double xm = 0, ym = 0, zm = 0, rm = 0;
        bool mine = false;
        int pm_ = 0;
        if (inm) {
            mine = !a.done[tm];
            const double4 q = a.p4[tm];
            xm = q.x; ym = q.y; zm = q.z; rm = a.r[tm];
            pm_ = a.passes[tm];
        }
        const uint64_t todo = __ballot(mine);
        // ranges still to serve: a stack of bit masks of members in LDS,
        // depth first (the group splits into at most 8 chunks, a chunk into
        // halves: at most 8 + 3 entries)
        int rs = 0;
        if (todo != 0) {
            if (threadIdx.x == 0) {
                rstack[0] = todo;
                rmode[0] = Shape::UNION_FIRST;
            }
            rs = 1;
        }
        __syncthreads();
        int handoff = 0, again = 0;
        while (rs > 0) {
            --rs;
            const uint64_t rg = rstack[rs];
            const int um = rmode[rs];
            const int nr = __popcll(rg);
            const bool in_r = (rg >> lane) & 1;
            const Sphere S = wave_sphere(in_r, xm, ym, zm, rm);
            int nc = -1;
            if (um == 1 || nr == 1 || S.R <= GROUP_REACH * S.rmax) {
                // member spheres of the range (widened like R) in the LDS of
                // the histograms, free until the selections
                double4* msph = reinterpret_cast<double4*>(&hist[0][0]);
                const int nm = (um == 1 && nr > 1) ? nr : 0;
                if (nm > 0 && w == 0 && in_r) {
                    const double rr = rm * (1.0 + 1e-12) + 1e-300;
                    msph[__popcll(rg & gf_lanemask_lt())] = make_double4(xm, ym, zm, rr * rr);
                }
                __syncthreads();
                const int root =
                    nr == 1 ? containing_node(a, a.leaf_of[b0 + __ffsll((unsigned long long)rg) - 1],
                                              S.bx, S.by, S.bz, S.R)
                            : group_containing_node(a, gp, S.bx, S.by, S.bz, S.R);
                nc = gather_group<CMAX>(a, root, S.bx, S.by, S.bz, S.R * S.R, msph, nm, cx, cy, cz, ci,
                                  cap, stack, lf, lincl, ctl);
            }
            if (nc == -2) {
                handoff = 1;
            } else if (nc >= 0) {
                // each wave selects for every fourth member of the range
                int rank_in_r = 0;
                for (int mm = 0; mm < np; ++mm) {
                    if (((rg >> mm) & 1) == 0) continue;
                    if ((rank_in_r++ & 3) != w) continue;
                    const double xi = gf_readlane_d(xm, mm), yi = gf_readlane_d(ym, mm);
                    const double zi = gf_readlane_d(zm, mm), ri = gf_readlane_d(rm, mm);
                    const int npass = __builtin_amdgcn_readlane(pm_, mm);
                    const int res = select_member<CMAX>(a, b0 + mm, xi, yi, zi, ri, npass, cx, cy, cz,
                                                  ci, nc, hist[w], tie_d[w], tie_i[w], tie_p[w]);
                    if (res < 0) handoff = 1;
                    if (res == 0) again = 1;
                }
            } else if (nr > 1 && um == 0) {
                // the bounding sphere is too large: gather the union of the
                // members' own spheres instead
                __syncthreads();              // every wave has read rstack[rs]
                if (threadIdx.x == 0) {
                    rstack[rs] = rg;
                    rmode[rs] = 1;
                }
                rs += 1;
            } else if (nr > 1) {
                // split the range into C chunks of consecutive members (unions
                // too): as many as its count asks for (2..8) at the group
                // level, halves below
                int C = 2;
                if (rg == todo && ctl[7] > 0) {
                    const long long want = (2ll * ctl[7] + cap - 1) / cap;
                    while (C < want && C < 8) C <<= 1;
                }
                C = min(C, nr);
                __syncthreads();              // every wave has read rstack[rs] and ctl[7]
                if (rs + C > 16) {
                    // cannot happen (at most 8 + 3 entries); an error, not a write
                    if (threadIdx.x == 0) atomicOr(&a.state[13], 2);
                    handoff = 1;
                    rs = 0;
                } else {
                    if (threadIdx.x == 0) {
                        uint64_t rest = rg;
                        for (int c = 0; c < C; ++c) {
                            // chunk c takes the next (nr - taken) / (C - c) members
                            const int take = (__popcll(rest) + (C - c) - 1) / (C - c);
                            uint64_t part = 0;
                            for (int q = 0; q < take; ++q) {
                                const uint64_t bit = rest & (~rest + 1);
                                part |= bit;
                                rest ^= bit;
                            }
                            rstack[rs + C - 1 - c] = part;    // first chunk on top
                            rmode[rs + C - 1 - c] = 1;
                        }
                    }
                    rs += C;
                }
            } else {
                // one member whose sphere holds more than the list (ctl[7]
                // particles at least, 0 = unknown): shrink the radius towards
                // 1.3 kq inside, or bisect inside the bracket; a bracket
                // narrower than 5 % (the count jumps past the list) goes to
                // the heap pass
                const int mm = __ffsll((unsigned long long)rg) - 1;
                const int t = b0 + mm;
                const double ri = gf_readlane_d(rm, mm);
                const int npass = __builtin_amdgcn_readlane(pm_, mm);
                const int seen = ctl[7];
                if (threadIdx.x == 0) {
                    const double lo = a.r_lo[t];
                    const bool closed = lo > 0 && ri <= lo * 1.05;
                    a.passes[t] = npass + 1;
                    a.r_hi[t] = ri;
                    double rn = seen > 0 ? ri * fmin(0.9, fmax(0.3, cbrt(1.3 * a.kq / double(seen))))
                                         : 0.5 * ri;
                    if (rn <= lo) rn = sqrt(lo * ri);
                    if (!closed) {
                        a.r[t] = rn;
                        atomicAdd(&a.state[0], 1);
                        ctl[3] |= 2;
                    } else {
                        ctl[3] |= 1;
                    }
                }
            }
            // the list is reused by the next range
            __syncthreads();
        }
        if (lane == 0 && handoff) atomicOr(&ctl[3], 1);
        if (lane == 0 && again) atomicOr(&ctl[3], 2);
        __syncthreads();
        if (threadIdx.x == 0) {
            // the heap pass and the next pass take the work's groups (members
            // already resolved are skipped there)
            for (int g = 0; g < ngrp; ++g) {
                if (ctl[3] & 1) a.ovf[atomicAdd(&a.state[3], 1)] = gp + g;
                if (ctl[3] & 2) request_retry(a, gp + g);
            }
            ctl[3] = 0;
        }
        __syncthreads();
    }
}

// Sorts each row of a solved neighbour table by (d^2, input index),
// ascending, the order the heap pass writes in its sorted mode (sph.py:165-166
// order): one wave a row, its kq <= 64 NS entries on NS slots a lane, a
// bitonic network of shuffles.  d^2 is the selection's FP64 value (member
// minus neighbour, no contraction), so the order is that of the selection.
template <int NS>
__global__ __launch_bounds__(256) void k_sort_rows(int n, int kq, const double4* __restrict__ p4,
                                                   const uint32_t* __restrict__ idx,
                                                   const int* __restrict__ inv, int tree_pos_out,
                                                   int* __restrict__ nb) {
#pragma clang fp contract(off)
    const int lane = __lane_id();
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= n) return;                    // whole waves leave together
    const double4 q = p4[t];
    double d[NS];
    int id[NS], pos[NS];
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const int sl = lane + 64 * r;
        d[r] = INFINITY;
        id[r] = 0x7fffffff;
        pos[r] = 0;
        if (sl < kq) {
            const int v = nb[size_t(sl) * n + t];
            const int ps = tree_pos_out ? v : inv[v];
            const double4 c = p4[ps];
            const double dx = q.x - c.x, dy = q.y - c.y, dz = q.z - c.z;
            d[r] = dx * dx + dy * dy + dz * dz;
            id[r] = int(idx[ps]);
            pos[r] = ps;
        }
    }
    constexpr int N = 64 * NS;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j == 64) {
                // NS == 2, k == 128: slot 0 keeps the smaller of the lane's two
                if constexpr (NS == 2) {
                    if (key_less(d[1], id[1], d[0], id[0])) {
                        const double td = d[0]; d[0] = d[1]; d[1] = td;
                        const int ti = id[0]; id[0] = id[1]; id[1] = ti;
                        const int tp = pos[0]; pos[0] = pos[1]; pos[1] = tp;
                    }
                }
                continue;
            }
#pragma unroll
            for (int r = 0; r < NS; ++r) {
                const int i = lane + 64 * r;
                const double od = __shfl_xor(d[r], j, 64);
                const int oi = __shfl_xor(id[r], j, 64);
                const int op = __shfl_xor(pos[r], j, 64);
                const bool want_min = ((i & j) == 0) == ((i & k) == 0);
                const bool take = want_min ? key_less(od, oi, d[r], id[r])
                                           : key_less(d[r], id[r], od, oi);
                if (take) {
                    d[r] = od;
                    id[r] = oi;
                    pos[r] = op;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const int sl = lane + 64 * r;
        if (sl < kq) nb[size_t(sl) * n + t] = tree_pos_out ? pos[r] : id[r];
    }
}

// Packed gas-tree nodes and particle positions for the neighbour walks.
__global__ __launch_bounds__(TB) void k_knode_pack(const int* __restrict__ counters,
                                                   const double4* __restrict__ box,
                                                   const int8_t* __restrict__ leaf,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ count,
                                                   const int* __restrict__ child8,
                                                   KNode* __restrict__ kn) {
    const int nn = counters[0];
    for (int id = blockIdx.x * TB + threadIdx.x; id < nn; id += gridDim.x * TB) {
        KNode k;
        k.box = box[id];
        if (leaf[id] || count[id] <= KNN_RANGE) {
            k.lo = make_int4(KNODE_RANGE, start[id], count[id],
                             leaf[id] && count[id] > GF_LEAF_SIZE ? 1 : 0);
            k.hi = make_int4(-1, -1, -1, -1);
        } else {
            const int4* cs = reinterpret_cast<const int4*>(child8 + 8 * size_t(id));
            k.lo = cs[0];
            k.hi = cs[1];
        }
        kn[id] = k;
    }
}

__global__ __launch_bounds__(TB) void k_p4_pack(int n, const double* __restrict__ tx,
                                                const double* __restrict__ ty,
                                                const double* __restrict__ tz,
                                                double4* __restrict__ p4) {
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t < n) p4[t] = make_double4(tx[t], ty[t], tz[t], 0.0);
}

// First search radius: the caller's, or from the number density of the
// smallest tree cell around the particle that holds >= kq particles.
__global__ __launch_bounds__(TB) void k_knn_init(int n, int kq, const double* __restrict__ r_init,
                                                 const uint32_t* __restrict__ idx,
                                                 const int* __restrict__ leaf_of,
                                                 const int* __restrict__ parent,
                                                 const int* __restrict__ count,
                                                 const double4* __restrict__ box,
                                                 double* __restrict__ r, double* __restrict__ r_lo,
                                                 double* __restrict__ r_hi,
                                                 int8_t* __restrict__ done,
                                                 int* __restrict__ passes,
                                                 int* __restrict__ gmark,
                                                 int* __restrict__ state) {
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= n) return;
    if (t == 0) state[14] = 0;      // merged work items (k_knn_items counts them)
    gmark[t] = 0;                   // groups <= particles
    double rr;
    if (r_init) {
        rr = r_init[idx[t]];
    } else {
        int node = leaf_of[t];
        while (count[node] < kq && parent[node] >= 0) node = parent[node];
        const double side = 2.0 * box[node].w;
        const double dens = double(count[node]) / (side * side * side);
        rr = 1.1 * cbrt(3.0 * kq / (4.0 * M_PI * dens));
    }
    r[t] = rr;
    r_lo[t] = 0.0;
    r_hi[t] = INFINITY;
    done[t] = 0;
    passes[t] = 0;
}

// The node of every group and its ancestors, GF_ANC per group (-1 past the root).
__global__ __launch_bounds__(TB) void k_group_anc(const int* __restrict__ counters,
                                                  const int* __restrict__ groups,
                                                  const int* __restrict__ parent,
                                                  int* __restrict__ anc) {
    const int ng = counters[1];
    for (int g = blockIdx.x * TB + threadIdx.x; g < ng; g += gridDim.x * TB) {
        int node = groups[g];
        for (int k = 0; k < GF_ANC; ++k) {
            anc[size_t(g) * GF_ANC + k] = node;
            node = node >= 0 ? parent[node] : -1;
        }
    }
}

__global__ __launch_bounds__(TB) void k_leaf_of(const int* __restrict__ counters,
                                                const int* __restrict__ start,
                                                const int* __restrict__ count,
                                                const int8_t* __restrict__ leaf,
                                                int* __restrict__ leaf_of) {
    const int nn = counters[0];
    for (int id = blockIdx.x * TB + threadIdx.x; id < nn; id += gridDim.x * TB) {
        if (!leaf[id]) continue;
        for (int i = start[id], e = start[id] + count[id]; i < e; ++i) leaf_of[i] = id;
    }
}

__global__ __launch_bounds__(TB) void k_sph_gather(int n, const uint32_t* __restrict__ idx,
                                                   const double* __restrict__ vel,
                                                   const double* __restrict__ u,
                                                   double* __restrict__ vx, double* __restrict__ vy,
                                                   double* __restrict__ vz, double* __restrict__ uu,
                                                   int* __restrict__ inv) {
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= n) return;
    const int j = int(idx[t]);
    inv[j] = t;
    if (vel) {
        vx[t] = vel[3 * j];
        vy[t] = vel[3 * j + 1];
        vz[t] = vel[3 * j + 2];
    }
    if (u) uu[t] = u[j];
}

// -------------------------------------------------------- density / forces
// sph.py:31-40 and :43-54 with the per-particle factors 1/h, 1/(pi h^3) and
// 1/(pi h^4) computed once per particle.
__device__ __forceinline__ double kernel_w(double r, double ihv, double s3) {
#pragma clang fp contract(off)
    const double q = r * ihv;
    if (q < 1.0) return s3 * (1 - 1.5 * q * q + 0.75 * (q * q * q));
    if (q < 2.0) {
        const double t = 2 - q;
        return s3 * 0.25 * (t * t * t);
    }
    return 0.0;
}

__device__ __forceinline__ double kernel_g(double r, double ihv, double s4) {
#pragma clang fp contract(off)
    const double q = r * ihv;
    double g;
    if (q < 1.0) {
        g = s4 * (-3 * q + 2.25 * q * q);
    } else if (q < 2.0) {
        const double t = 2 - q;
        g = -s4 * 0.75 * (t * t);
    } else {
        return 0.0;
    }
    return g / fmax(r, 1e-30);
}

struct SphView {
    const double *x, *y, *z, *vx, *vy, *vz, *m, *u, *h;
    int n;
};

// What a gas particle contributes as a neighbour, packed in 128 bytes so one
// neighbour read touches two cache lines instead of one per array:
//   a = (x, y, z, m), b = (vx, vy, vz, c), c = (h, 1/h, 1/(pi h^4), rho),
//   d = (P / rho^2, -, -, -),  with c = sqrt(gamma (gamma - 1) u) and
//   P = (gamma - 1) rho u: the expressions of sph.py:73-74, :88-89, :98.
struct SphRec {
    double4 a, b, c, d;
};

__global__ __launch_bounds__(TB) void k_pack(SphView v, double gamma, SphRec* __restrict__ rec,
                                             double4* __restrict__ a4) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= v.n) return;
    const double hv = v.h[t];
    SphRec r;
    r.a = make_double4(v.x[t], v.y[t], v.z[t], v.m[t]);
    a4[t] = r.a;
    r.b = make_double4(v.vx[t], v.vy[t], v.vz[t], sqrt(gamma * (gamma - 1) * v.u[t]));
    r.c = make_double4(hv, 1.0 / hv, 1.0 / (M_PI * (hv * hv * hv * hv)), 0.0);
    r.d = make_double4(0.0, 0.0, 0.0, 0.0);
    rec[t] = r;
}

// rho_i = sum_j m_j W(r_ij, h_i) over the neighbour table (sph.py:63-71),
// neighbours read from the compact (x, y, z, m) copy (four per cache line);
// also completes the particle's record with rho and P / rho^2.
// Lanes per particle of the density and force sums: lane q of a particle sums
// its neighbours q, q + SPH_LANES, ...; the partial sums combine by
// butterfly steps.  With four lanes per particle and the blocks of each XCD
// in particle order, an XCD has about 16k particles in flight, whose
// neighbours' 128-byte records (about 2 MB) fit its 4 MB L2.
#ifndef SPH_LANES
#define SPH_LANES 4
#endif
constexpr int SL = SPH_LANES;

template <typename T>
__device__ __forceinline__ T lanes_sum(T v) {
#pragma unroll
    for (int o = 1; o < SL; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double lanes_max(double v) {
#pragma unroll
    for (int o = 1; o < SL; o <<= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ __launch_bounds__(TB) void k_density(int n, const int* __restrict__ nb, int k,
                                                const double* __restrict__ u, double gamma,
                                                const double4* __restrict__ a4,
                                                SphRec* __restrict__ rec,
                                                double* __restrict__ rho) {
#pragma clang fp contract(off)
    const int gt = gf_xcd_block() * TB + threadIdx.x;   // particles in order per XCD
    const int t = gt / SL, q = gt % SL;
    const bool ok = t < n;                    // every lane takes part in the sums
    const int tt = ok ? t : 0;
    const double4 at = a4[tt];
    const double4 ct = rec[tt].c;
    const double hv = ct.x, ihv = ct.y, s3 = 1.0 / (M_PI * (hv * hv * hv));
    double s = 0.0;
    if (ok) {
        for (int jj = q; jj < k; jj += SL) {
            const int j = nb[size_t(jj) * n + t];
            const double4 aj = a4[j];
            const double dx = at.x - aj.x, dy = at.y - aj.y, dz = at.z - aj.z;
            s += aj.w * kernel_w(sqrt(dx * dx + dy * dy + dz * dz), ihv, s3);
        }
    }
    s = lanes_sum(s);
    if (!ok || q != 0) return;
    rho[t] = s;
    rec[t].c.w = s;
    const double P = (gamma - 1) * s * u[t];
    rec[t].d.x = P / (s * s);
}

// Pressure + viscosity acceleration, du/dt, Courant step (sph.py:72-106) and
// div v = -(1/rho_i) sum_j m_j (v_ij . r_ij) (1/r) dW/dr(h_i).
__global__ __launch_bounds__(TB) void k_force(int n, const int* __restrict__ nb, int k,
                                              const SphRec* __restrict__ rec, double alpha,
                                              double beta, const uint32_t* __restrict__ out_idx,
                                              double* __restrict__ acc, double* __restrict__ dudt,
                                              double* __restrict__ dt, double* __restrict__ divv,
                                              double* __restrict__ rho_out,
                                              double* __restrict__ h_out) {
#pragma clang fp contract(off)
    const int gt = gf_xcd_block() * TB + threadIdx.x;   // particles in order per XCD
    const int i = gt / SL, q = gt % SL;
    const bool ok = i < n;
    const SphRec ri = rec[ok ? i : 0];
    const double xi = ri.a.x, yi = ri.a.y, zi = ri.a.z;
    const double vxi = ri.b.x, vyi = ri.b.y, vzi = ri.b.z, ci = ri.b.w;
    const double hi_ = ri.c.x, ihi = ri.c.y, sgi = ri.c.z, rhoi = ri.c.w;
    const double pi2 = ri.d.x;
    double ax = 0, ay = 0, az = 0, du = 0, mu_max = 0, dv = 0;
    if (ok) {
        for (int jj = q; jj < k; jj += SL) {
            const int j = nb[size_t(jj) * n + i];
            if (j == i) continue;
            const double4 aj = rec[j].a, bj = rec[j].b, cj4 = rec[j].c;
            const double pj2 = rec[j].d.x;
            const double dx = xi - aj.x, dy = yi - aj.y, dz = zi - aj.z;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const double gi = kernel_g(r, ihi, sgi);
            const double gj = kernel_g(r, cj4.y, cj4.z);
            const double rhoj = cj4.w, hj = cj4.x, cj = bj.w;
            const double vr = (vxi - bj.x) * dx + (vyi - bj.y) * dy + (vzi - bj.z) * dz;
            double visc = 0.0;
            if (vr < 0) {
                const double hb = 0.5 * (hi_ + hj);
                const double mu = hb * vr / (r * r + 0.01 * hb * hb);
                mu_max = fmax(mu_max, -mu);
                visc = (-alpha * 0.5 * (ci + cj) * mu + beta * mu * mu) / (0.5 * (rhoi + rhoj));
            }
            const double gb = 0.5 * (gi + gj);
            const double f = pi2 * gi + pj2 * gj + visc * gb;
            const double mj = aj.w;
            ax -= mj * f * dx;
            ay -= mj * f * dy;
            az -= mj * f * dz;
            du += mj * (pi2 * gi + 0.5 * visc * gb) * vr;
            dv += mj * vr * gi;
        }
    }
    ax = lanes_sum(ax);
    ay = lanes_sum(ay);
    az = lanes_sum(az);
    du = lanes_sum(du);
    dv = lanes_sum(dv);
    mu_max = lanes_max(mu_max);
    if (!ok || q != 0) return;
    const size_t o = out_idx ? size_t(out_idx[i]) : size_t(i);
    if (acc) {
        acc[3 * o] = ax;
        acc[3 * o + 1] = ay;
        acc[3 * o + 2] = az;
    }
    if (dudt) dudt[o] = du;
    if (dt) dt[o] = 0.3 * hi_ / (ci + 1.2 * (alpha * ci + beta * mu_max) + 1e-30);
    if (divv) divv[o] = -dv / rhoi;
    if (rho_out) rho_out[o] = rhoi;
    if (h_out) h_out[o] = hi_;
}

// FP32 pair arithmetic (the fast path's density and forces): separations
// are formed in FP64 from the FP64 positions and rounded to FP32 once, the
// pair terms are FP32 (v_sqrt_f32, v_rcp_f32, no branches), and every sum is
// accumulated in FP64.  The neighbour table, h and the order of the sums are
// those of the FP64 kernels; the density differs from theirs by FP32
// rounding of the terms (about 1e-7 relative).
// Record per particle (48 bytes, in the FP64 record's memory): v = (vx, vy,
// vz, c), k = (h, 1/h, 1/(pi h^4), rho), d = (P / rho^2, m, 1/(pi h^3), -).
struct SphRec32 {
    float4 v, k, d;
};

__device__ __forceinline__ float kernel_w32(float r, float ihv, float s3) {
    const float q = r * ihv;
    const float t = fmaxf(2.0f - q, 0.0f);
    const float inner = 1.0f - 1.5f * q * q + 0.75f * (q * q * q);
    return s3 * (q < 1.0f ? inner : 0.25f * (t * t * t));
}

// (1/r) dW/dr without the 1/r: s4 {-3q + 2.25q^2; -0.75(2 - q)^2; 0}
__device__ __forceinline__ float kernel_g32(float r, float ihv, float s4) {
    const float q = r * ihv;
    const float t = fmaxf(2.0f - q, 0.0f);
    return s4 * (q < 1.0f ? (-3.0f * q + 2.25f * q * q) : -0.75f * (t * t));
}

__global__ __launch_bounds__(TB) void k_pack32(SphView v, double gamma, SphRec32* __restrict__ rec,
                                               double4* __restrict__ a4) {
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= v.n) return;
    const double hv = v.h[t];
    a4[t] = make_double4(v.x[t], v.y[t], v.z[t], v.m[t]);
    SphRec32 r;
    r.v = make_float4(float(v.vx[t]), float(v.vy[t]), float(v.vz[t]),
                      float(sqrt(gamma * (gamma - 1) * v.u[t])));
    r.k = make_float4(float(hv), float(1.0 / hv), float(1.0 / (M_PI * (hv * hv * hv * hv))), 0.0f);
    r.d = make_float4(0.0f, float(v.m[t]), float(1.0 / (M_PI * (hv * hv * hv))), 0.0f);
    rec[t] = r;
}

__global__ __launch_bounds__(TB) void k_density32(int n, const int* __restrict__ nb, int k,
                                                  const double* __restrict__ u, double gamma,
                                                  const double4* __restrict__ a4,
                                                  SphRec32* __restrict__ rec,
                                                  double* __restrict__ rho) {
    const int gt = gf_xcd_block() * TB + threadIdx.x;   // particles in order per XCD
    const int t = gt / SL, q = gt % SL;
    const bool ok = t < n;
    const int tt = ok ? t : 0;
    const double4 at = a4[tt];
    const float ihv = rec[tt].k.y, s3 = rec[tt].d.z;
    double s = 0.0;
    if (ok) {
        for (int jj = q; jj < k; jj += SL) {
            const int j = nb[size_t(jj) * n + t];
            const double4 aj = a4[j];
            const float dx = float(at.x - aj.x), dy = float(at.y - aj.y), dz = float(at.z - aj.z);
            const float r = __builtin_sqrtf(dx * dx + dy * dy + dz * dz);
            s += double(float(aj.w) * kernel_w32(r, ihv, s3));
        }
    }
    s = lanes_sum(s);
    if (!ok || q != 0) return;
    rho[t] = s;
    rec[t].k.w = float(s);
    const double P = (gamma - 1) * s * u[t];
    rec[t].d.x = float(P / (s * s));
}

__global__ __launch_bounds__(TB) void k_force32(int n, const int* __restrict__ nb, int k,
                                                const double4* __restrict__ a4,
                                                const SphRec32* __restrict__ rec,
                                                const double* __restrict__ rho_d,
                                                const double* __restrict__ h_d, double alpha_d,
                                                double beta_d, const uint32_t* __restrict__ out_idx,
                                                double* __restrict__ acc, double* __restrict__ dudt,
                                                double* __restrict__ dt, double* __restrict__ divv,
                                                double* __restrict__ rho_out,
                                                double* __restrict__ h_out) {
    const int gt = gf_xcd_block() * TB + threadIdx.x;   // particles in order per XCD
    const int i = gt / SL, q = gt % SL;
    const bool ok = i < n;
    const int ii = ok ? i : 0;
    const double4 ai = a4[ii];
    const SphRec32 ri = rec[ii];
    const float alpha = float(alpha_d), beta = float(beta_d);
    const float vxi = ri.v.x, vyi = ri.v.y, vzi = ri.v.z, ci = ri.v.w;
    const float hi_ = ri.k.x, ihi = ri.k.y, sgi = ri.k.z, rhoi = ri.k.w;
    const float pi2 = ri.d.x;
    double ax = 0, ay = 0, az = 0, du = 0, dv = 0;
    float mu_max = 0.0f;
    if (ok) {
        for (int jj = q; jj < k; jj += SL) {
            const int j = nb[size_t(jj) * n + i];
            if (j == i) continue;
            const double4 aj = a4[j];
            const SphRec32 rj = rec[j];
            const float dx = float(ai.x - aj.x), dy = float(ai.y - aj.y), dz = float(ai.z - aj.z);
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float r = __builtin_sqrtf(r2);
            const float ir = __builtin_amdgcn_rcpf(fmaxf(r, 1e-30f));
            const float gi = kernel_g32(r, ihi, sgi) * ir;
            const float gj = kernel_g32(r, rj.k.y, rj.k.z) * ir;
            const float vr = (vxi - rj.v.x) * dx + (vyi - rj.v.y) * dy + (vzi - rj.v.z) * dz;
            const float hb = 0.5f * (hi_ + rj.k.x);
            const float mu = hb * vr * __builtin_amdgcn_rcpf(r2 + 0.01f * hb * hb);
            const float vis = (-alpha * 0.5f * (ci + rj.v.w) * mu + beta * mu * mu) *
                              (2.0f * __builtin_amdgcn_rcpf(rhoi + rj.k.w));
            const bool app = vr < 0.0f;             // viscosity only for approaching pairs
            const float visc = app ? vis : 0.0f;
            mu_max = app ? fmaxf(mu_max, -mu) : mu_max;
            const float gb = 0.5f * (gi + gj);
            const float f = pi2 * gi + rj.d.x * gj + visc * gb;
            const float mj = rj.d.y;
            const float mf = mj * f;
            ax -= double(mf * dx);
            ay -= double(mf * dy);
            az -= double(mf * dz);
            du += double(mj * (pi2 * gi + 0.5f * visc * gb) * vr);
            dv += double(mj * vr * gi);
        }
    }
    ax = lanes_sum(ax);
    ay = lanes_sum(ay);
    az = lanes_sum(az);
    du = lanes_sum(du);
    dv = lanes_sum(dv);
    const double mm = lanes_max(double(mu_max));
    if (!ok || q != 0) return;
    const size_t o = out_idx ? size_t(out_idx[i]) : size_t(i);
    const double rho_i = rho_d[i], hd = h_d[i], cd = double(ci);
    if (acc) {
        acc[3 * o] = ax;
        acc[3 * o + 1] = ay;
        acc[3 * o + 2] = az;
    }
    if (dudt) dudt[o] = du;
    if (dt) dt[o] = 0.3 * hd / (cd + 1.2 * (alpha_d * cd + beta_d * mm) + 1e-30);
    if (divv) divv[o] = -dv / rho_i;
    if (rho_out) rho_out[o] = rho_i;
    if (h_out) h_out[o] = hd;
}

__global__ __launch_bounds__(TB) void k_smoothing(int n, const double* __restrict__ dk, double eps,
                                                  double* __restrict__ h) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= n) return;
    h[t] = fmax(0.5 * dk[t], 0.5 * eps);
}

// caller's table: (n, k) int64 -> [k][n] int32, SoA copies of pos / vel
__global__ __launch_bounds__(TB) void k_table_prep(int n, int k, const int64_t* __restrict__ nb,
                                                   const double* __restrict__ pos,
                                                   const double* __restrict__ vel,
                                                   int* __restrict__ nbt, double* __restrict__ soa) {
    const int i = blockIdx.x * TB + threadIdx.x;
    if (i >= n) return;
    for (int jj = 0; jj < k; ++jj) nbt[size_t(jj) * n + i] = int(nb[size_t(i) * k + jj]);
    for (int c = 0; c < 3; ++c) {
        soa[size_t(c) * n + i] = pos[3 * size_t(i) + c];
        soa[size_t(3 + c) * n + i] = vel[3 * size_t(i) + c];
    }
}

__global__ __launch_bounds__(TB) void k_export_knn(int n, int kq, const int* __restrict__ nb,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z,
                                                   const uint32_t* __restrict__ idx,
                                                   const int* __restrict__ inv,
                                                   int64_t* __restrict__ onb,
                                                   double* __restrict__ odist) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= n) return;
    const size_t row = size_t(idx[t]) * kq;
    for (int s = 0; s < kq; ++s) {
        const int id = nb[size_t(s) * n + t];
        onb[row + s] = id;
        const int j = inv[id];
        const double dx = x[t] - x[j], dy = y[t] - y[j], dz = z[t] - z[j];
        odist[row + s] = sqrt(dx * dx + dy * dy + dz * dz);
    }
}

// sph.py:109-135: k nearest of each row's candidates, ascending (d^2, index).
constexpr int CAND_MAX = 128;
__global__ __launch_bounds__(64) void k_knn_cand(int n, int m, const double* __restrict__ pos,
                                                 const int64_t* __restrict__ cand, int k,
                                                 int64_t* __restrict__ nb, double* __restrict__ dist,
                                                 int* __restrict__ ok) {
#pragma clang fp contract(off)
    __shared__ double d2s[CAND_MAX * 64];
    __shared__ int ids[CAND_MAX * 64];
    const int lane = threadIdx.x;
    const int i = blockIdx.x * 64 + lane;
    if (i >= n) return;
    const double x = pos[3 * size_t(i)], y = pos[3 * size_t(i) + 1], z = pos[3 * size_t(i) + 2];
    int cnt = 0;
    for (int jj = 0; jj < m; ++jj) {
        const int64_t j = cand[size_t(i) * m + jj];
        if (j < 0) continue;
        const double dx = x - pos[3 * j], dy = y - pos[3 * j + 1], dz = z - pos[3 * j + 2];
        d2s[cnt * 64 + lane] = dx * dx + dy * dy + dz * dz;
        ids[cnt * 64 + lane] = int(j);
        ++cnt;
    }
    if (cnt < k) {
        atomicAnd(ok, 0);
        return;
    }
    for (int a0 = 0; a0 < cnt; ++a0) {
        const double da = d2s[a0 * 64 + lane];
        const int ia = ids[a0 * 64 + lane];
        // equal keys (a candidate listed twice) keep their column order, as
        // a stable sort would
        int rank = 0;
        for (int b = 0; b < cnt; ++b) {
            const double db = d2s[b * 64 + lane];
            const int ib = ids[b * 64 + lane];
            rank += key_less(db, ib, da, ia) || (b < a0 && db == da && ib == ia);
        }
        if (rank < k) {
            nb[size_t(i) * k + rank] = ia;
            dist[size_t(i) * k + rank] = sqrt(da);
        }
    }
}

__global__ __launch_bounds__(TB) void k_kernel_eval(int n, const double* __restrict__ r,
                                                    const double* __restrict__ h,
                                                    double* __restrict__ w,
                                                    double* __restrict__ g) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= n) return;
    const double hv = h[t], rv = r[t];
    // the reference's own operation order (q = r / h), sph.py:34-35, :46-47
    const double q = rv / hv;
    if (w) {
        const double s = 1.0 / (M_PI * (hv * hv * hv));
        double out = 0.0;
        if (q < 1.0) out = s * (1 - 1.5 * q * q + 0.75 * (q * q * q));
        else if (q < 2.0) { const double tt = 2 - q; out = s * 0.25 * (tt * tt * tt); }
        w[t] = out;
    }
    if (g) {
        const double s = 1.0 / (M_PI * (hv * hv * hv * hv));
        double gg = 0.0;
        if (q < 1.0) gg = s * (-3 * q + 2.25 * q * q) / fmax(rv, 1e-30);
        else if (q < 2.0) { const double tt = 2 - q; gg = -s * 0.75 * (tt * tt) / fmax(rv, 1e-30); }
        g[t] = gg;
    }
}

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

int sph_reserve(GfSph& S, int n, int k) {
    if (!S.state) {
        if (dalloc(S.state, 16) || dalloc(S.gstack, size_t(KNN_GST_BLOCKS) * KGSCAP)) return -1;
    }
    if (n > S.cap_n) {
        const int cn = std::max(n, 1024);
        if (dalloc(S.vx, cn) || dalloc(S.vy, cn) || dalloc(S.vz, cn) || dalloc(S.u, cn) ||
            dalloc(S.r, cn) || dalloc(S.dk, cn) || dalloc(S.h, cn) || dalloc(S.rho, cn) ||
            dalloc(S.done, cn) || dalloc(S.passes, cn) ||
            dalloc(S.inv, cn) || dalloc(S.leaf_of, cn) ||
            dalloc(S.r_lo, cn) || dalloc(S.r_hi, cn) || dalloc(S.rec, 16 * size_t(cn)) ||
            dalloc(S.p4, 4 * size_t(cn)) || dalloc(S.gmark, cn))
            return -1;
        S.cap_n = cn;
        S.cap_k = 0;
    }
    if (k > S.cap_k) {
        if (dalloc(S.nb, size_t(S.cap_n) * size_t(k))) return -1;
        S.cap_k = k;
    }
    const int ng = gf_tree_node_bound(S.cap_n);
    if (ng > S.cap_groups) {
        if (dalloc(S.ovf, ng) || dalloc(S.ovf2, ng) || dalloc(S.retry, 2 * size_t(ng)) ||
            dalloc(S.knode, 8 * size_t(ng)) || dalloc(S.anc, size_t(GF_ANC) * size_t(ng)) ||
            dalloc(S.items, 4 * size_t(ng)))
            return -1;
        S.cap_groups = ng;
    }
    return 0;
}

}  // namespace

int gf_sph_fast_blocks(const GfSph& S) { return S.fast_blocks > 0 ? S.fast_blocks : FAST_BLOCKS; }

int gf_sph_merge_members(const GfSph& S) {
    return S.merge_members > 0 ? S.merge_members : KNN_MERGE_MEMBERS;
}

static int wide_blocks(const GfSph& S) { return S.fast_blocks > 0 ? S.fast_blocks : WIDE_BLOCKS; }

void gf_sph_free(GfSph& S) {
    gf_tree_free(S.tree);
    dfree(S.vx); dfree(S.vy); dfree(S.vz); dfree(S.u); dfree(S.r); dfree(S.dk); dfree(S.h);
    dfree(S.rho); dfree(S.done); dfree(S.passes); dfree(S.inv); dfree(S.leaf_of); dfree(S.r_lo); dfree(S.r_hi); dfree(S.rec); dfree(S.p4); dfree(S.gmark); dfree(S.retry); dfree(S.knode);
    dfree(S.anc); dfree(S.items);
    dfree(S.nb); dfree(S.state); dfree(S.ovf); dfree(S.ovf2); dfree(S.gstack);
    if (S.state_host) (void)hipHostFree(S.state_host);
    const int fb = S.fast_blocks, mm = S.merge_members;
    S = GfSph();
    S.fast_blocks = fb;
    S.merge_members = mm;
}

int gf_sph_setup(GfSph& S, int n, const double* pos, const double* vel, const double* mass,
                 const double* u, hipStream_t s) {
    if (sph_reserve(S, n, 0)) return -1;
    if (gf_tree_build(S.tree, n, pos, mass, 0.5f, s, /*with_com=*/false)) return -1;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_sph_gather, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, S.tree.idx, vel, u,
                       S.vx, S.vy, S.vz, S.u, S.inv);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_sph_setup_from_tree(GfSph& S, const GfTree& T, int n, const double* pos, const double* vel,
                           const double* mass, const double* u, hipStream_t s) {
    if (sph_reserve(S, n, 0)) return -1;
    if (gf_tree_build_subset(S.tree, T, n, pos, mass, s)) return -1;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_sph_gather, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, S.tree.idx, vel, u,
                       S.vx, S.vy, S.vz, S.u, S.inv);
    GF_LAUNCH_CHECK();
    return 0;
}

static KnnArgs knn_args(GfSph& S) {
    GfTree& T = S.tree;
    KnnArgs a;
    a.kn = reinterpret_cast<const KNode*>(S.knode);
    a.p4 = reinterpret_cast<const double4*>(S.p4);
    a.gmark = S.gmark;
    a.retry_in = nullptr;
    a.retry_out = nullptr;
    a.pass_tag = 0;
    a.list_cap = gf_test_limits().knn_list > 0 ? gf_test_limits().knn_list : CMAX_MAX;
    a.start = T.start; a.count = T.count; a.leaf = T.leaf; a.child8 = T.child8; a.box = T.box;
    a.groups = T.groups; a.counters = T.counters; a.tx = T.tx; a.ty = T.ty; a.tz = T.tz;
    a.idx = T.idx; a.inv = S.inv; a.parent_ = T.parent; a.leaf_of = S.leaf_of; a.r = S.r;
    a.anc = S.anc;
    a.r_lo = S.r_lo; a.r_hi = S.r_hi; a.dk = S.dk; a.done = S.done; a.passes = S.passes;
    a.nb = S.nb; a.n = T.n; a.kq = S.knn_kq; a.tree_pos_out = S.knn_pos_out;
    a.state = S.state; a.ovf = S.ovf; a.ovf2 = S.ovf2; a.gstack = S.gstack;
    a.items = S.knn_items ? reinterpret_cast<const int4*>(S.items) : nullptr;
    return a;
}

// Enqueues passes [from, from + count) of the held solve.  Each pass reads
// its work from the device (the previous pass's retry list), so a pass after
// convergence finds nothing to do and costs its launches only.
static int knn_enqueue(GfSph& S, int from, int count, hipStream_t s) {
    KnnArgs a = knn_args(S);
    int* rbuf[2] = {S.retry, S.retry + S.cap_groups};
    for (int pass = from; pass < from + count; ++pass) {
        // the previous pass's retry list is this pass's work
        a.retry_in = pass == 0 ? nullptr : rbuf[(pass - 1) & 1];
        a.retry_out = rbuf[pass & 1];
        a.pass_tag = pass + 1;
        hipLaunchKernelGGL(k_knn_reset, dim3(1), dim3(64), 0, s, S.state, pass == 0 ? 1 : 0);
        hipLaunchKernelGGL(S.knn_narrow ? k_knn_fast<KnnNarrow> : k_knn_fast<KnnWide>,
                           dim3(S.knn_narrow ? gf_sph_fast_blocks(S) : wide_blocks(S)),
                           dim3(FT), 0, s, a);
        if (S.knn_kq <= 64) {
            hipLaunchKernelGGL((k_knn_pass<64, false>), dim3(KNN_BLOCKS), dim3(64), 0, s, a);
            hipLaunchKernelGGL((k_knn_pass<64, true>), dim3(KNN_GST_BLOCKS), dim3(64), 0, s, a);
        } else {
            hipLaunchKernelGGL((k_knn_pass<128, false>), dim3(KNN_BLOCKS), dim3(64), 0, s, a);
            hipLaunchKernelGGL((k_knn_pass<128, true>), dim3(KNN_GST_BLOCKS), dim3(64), 0, s, a);
        }
        GF_LAUNCH_CHECK();
    }
    S.knn_enqueued = from + count;
    return 0;
}

int gf_sph_knn_begin(GfSph& S, int kq, const double* r_init, bool tree_pos_out, bool sorted,
                     int passes, hipStream_t s) {
    S.knn_open = 0;
    const int n = S.tree.n;
    if (n == 0) return 0;
    if (kq < 1 || kq > n) {
        gf_set_error("kNN: k = %d must be in [1, n = %d]", kq, n);
        return -1;
    }
    if (kq > 128) {
        gf_set_error("kNN: k = %d exceeds 128", kq);
        return -1;
    }
    if (passes < 0) {
        gf_set_error("kNN: a negative number of passes");
        return -1;
    }
    if (sph_reserve(S, n, kq)) return -1;
    GfTree& T = S.tree;
    // leaf of each particle: first radius and start of the member walks
    hipLaunchKernelGGL(k_leaf_of, dim3(std::min(2048, gf_div_up(T.cap_nodes, TB))), dim3(TB), 0, s,
                       T.counters, T.start, T.count, T.leaf, S.leaf_of);
    hipLaunchKernelGGL(k_group_anc, dim3(std::min(2048, gf_div_up(T.cap_nodes, TB))), dim3(TB), 0,
                       s, T.counters, T.groups, T.parent, S.anc);
    hipLaunchKernelGGL(k_knn_init, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, kq, r_init, T.idx,
                       S.leaf_of, T.parent, T.count, T.box, S.r, S.r_lo, S.r_hi, S.done,
                       S.passes, S.gmark, S.state);
    // from a caller's radii the first pass serves merged items (about k
    // candidates a member: sharing a gather pays); the cell-density start
    // overshoots, and its gathers stay per group
    S.knn_items = r_init && gf_sph_merge_members(S) > 1 ? 1 : 0;
    if (S.knn_items &&
        gf_tree_merge_groups(T, std::min(64, gf_sph_merge_members(S)), S.state + 14,
                             reinterpret_cast<int4*>(S.items), s))
        return -1;
    hipLaunchKernelGGL(k_knode_pack, dim3(std::min(2048, gf_div_up(T.cap_nodes, TB))), dim3(TB), 0,
                       s, T.counters, T.box, T.leaf, T.start, T.count, T.child8,
                       reinterpret_cast<KNode*>(S.knode));
    hipLaunchKernelGGL(k_p4_pack, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, T.tx, T.ty, T.tz,
                       reinterpret_cast<double4*>(S.p4));
    GF_LAUNCH_CHECK();
    S.knn_kq = kq;
    S.knn_pos_out = tree_pos_out ? 1 : 0;
    S.knn_sorted = sorted ? 1 : 0;
    S.knn_narrow = r_init ? 1 : 0;
    S.stat_queued = 0;
    S.stat_queued2 = 0;
    S.knn_enqueued = 0;
    if (passes > 0 && knn_enqueue(S, 0, passes, s)) return -1;
    S.knn_open = 1;
    return 0;
}

int gf_sph_knn_more(GfSph& S, int passes, hipStream_t s) {
    if (S.tree.n == 0) return 0;
    if (!S.knn_open) {
        gf_set_error("kNN: no solve was begun");
        return -1;
    }
    return passes > 0 ? knn_enqueue(S, S.knn_enqueued, passes, s) : 0;
}

int gf_sph_knn_end(GfSph& S, int max_passes, hipStream_t s) {
    if (S.tree.n == 0) return 0;
    if (!S.knn_open) {
        gf_set_error("kNN: no solve was begun");
        return -1;
    }
    S.knn_open = 0;
    // the counters come back through pinned memory: one short DMA after the
    // passes, where a pageable copy took tens of microseconds
    if (!S.state_host) GF_HIP(hipHostMalloc(&S.state_host, 16 * sizeof(int), hipHostMallocDefault));
    int* st = S.state_host;
    while (true) {
        GF_HIP(hipMemcpyAsync(st, S.state, 16 * sizeof(int), hipMemcpyDeviceToHost, s));
        GF_HIP(hipStreamSynchronize(s));
        if (st[4] != 0 || st[12] != 0) {
            gf_set_error("kNN: %d groups overflowed the global stack", st[4] + st[12]);
            return -1;
        }
        if (st[13] != 0) {
            gf_set_error("kNN: internal error %d (candidate range or member-range stack out of "
                         "bounds)", st[13]);
            return -1;
        }
        if (st[0] == 0 && S.knn_enqueued > 0) break;
        if (S.knn_enqueued >= max_passes) {
            gf_set_error("kNN: particles still unresolved after %d passes", max_passes);
            return -1;
        }
        // not converged: two more passes, then look again
        if (knn_enqueue(S, S.knn_enqueued, std::min(2, max_passes - S.knn_enqueued), s)) return -1;
    }
    S.stat_queued = st[9] + st[3];
    S.stat_queued2 = st[10] + st[6];
    S.last_passes = st[11];
    if (S.knn_sorted) {
        // the fast pass writes a row's neighbours unordered
        const GfTree& T = S.tree;
        const int nblk = gf_div_up(T.n, 4);
        if (S.knn_kq <= 64)
            hipLaunchKernelGGL(k_sort_rows<1>, dim3(nblk), dim3(256), 0, s, T.n, S.knn_kq,
                               reinterpret_cast<const double4*>(S.p4), T.idx, S.inv,
                               S.knn_pos_out, S.nb);
        else
            hipLaunchKernelGGL(k_sort_rows<2>, dim3(nblk), dim3(256), 0, s, T.n, S.knn_kq,
                               reinterpret_cast<const double4*>(S.p4), T.idx, S.inv,
                               S.knn_pos_out, S.nb);
        GF_LAUNCH_CHECK();
    }
    return 0;
}

int gf_sph_knn(GfSph& S, int kq, const double* r_init, bool tree_pos_out, bool sorted,
               int max_passes, hipStream_t s) {
    if (gf_sph_knn_begin(S, kq, r_init, tree_pos_out, sorted, std::min(2, max_passes), s))
        return -1;
    return gf_sph_knn_end(S, max_passes, s);
}

int gf_sph_smoothing(GfSph& S, int k, double eps, hipStream_t s) {
    (void)k;
    const int n = S.tree.n;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_smoothing, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, S.dk, eps, S.h);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_sph_forces_tree(GfSph& S, int k, double gamma, double alpha, double beta, double* rho,
                       double* acc, double* dudt, double* dt, double* divv, double* h_out,
                       hipStream_t s) {
    const int n = S.tree.n;
    if (n == 0) return 0;
    SphView v{S.tree.tx, S.tree.ty, S.tree.tz, S.vx, S.vy, S.vz, S.tree.tm, S.u, S.h, n};
    const int g = gf_div_up(n, TB);
    if (S.pair_f32) {
        static_assert(sizeof(SphRec32) <= 16 * sizeof(double), "record space");
        SphRec32* rec = reinterpret_cast<SphRec32*>(S.rec);
        double4* a4 = reinterpret_cast<double4*>(S.p4);  // free after the solve
        hipLaunchKernelGGL(k_pack32, dim3(g), dim3(TB), 0, s, v, gamma, rec, a4);
        const int gl = gf_div_up((long long)n * SL, TB);
        hipLaunchKernelGGL(k_density32, dim3(gl), dim3(TB), 0, s, n, S.nb, k, S.u, gamma, a4, rec,
                           S.rho);
        hipLaunchKernelGGL(k_force32, dim3(gl), dim3(TB), 0, s, n, S.nb, k, a4, rec, S.rho, S.h,
                           alpha, beta, S.tree.idx, acc, dudt, dt, divv, rho, h_out);
        GF_LAUNCH_CHECK();
        return 0;
    }
    SphRec* rec = reinterpret_cast<SphRec*>(S.rec);
    double4* a4 = reinterpret_cast<double4*>(S.p4);      // free after the solve
    hipLaunchKernelGGL(k_pack, dim3(g), dim3(TB), 0, s, v, gamma, rec, a4);
    const int gl = gf_div_up((long long)n * SL, TB);
    hipLaunchKernelGGL(k_density, dim3(gl), dim3(TB), 0, s, n, S.nb, k, S.u, gamma, a4, rec, S.rho);
    hipLaunchKernelGGL(k_force, dim3(gl), dim3(TB), 0, s, n, S.nb, k, rec, alpha, beta,
                       S.tree.idx, acc, dudt, dt, divv, rho, h_out);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_sph_forces_table(int n, int k, const double* pos, const double* vel, const double* mass,
                        const int64_t* nb, const double* h, const double* u, double gamma,
                        double alpha, double beta, double* rho, double* acc, double* dudt,
                        double* dt, double* divv, int* scratch, hipStream_t s) {
    if (n == 0) return 0;
    // scratch: [k][n] ints, then 6n doubles (SoA pos, vel), then n records
    int* nbt = scratch;
    double* soa = reinterpret_cast<double*>(scratch + ((size_t(k) * n + 1) & ~size_t(1)));
    SphRec* rec = reinterpret_cast<SphRec*>(
        (reinterpret_cast<uintptr_t>(soa + 6 * size_t(n)) + 255) & ~uintptr_t(255));
    double* rr = soa + 6 * size_t(n) + 16 * size_t(n) + 64;
    double4* a4 = reinterpret_cast<double4*>(
        (reinterpret_cast<uintptr_t>(rr + size_t(n)) + 255) & ~uintptr_t(255));
    const int g = gf_div_up(n, TB);
    hipLaunchKernelGGL(k_table_prep, dim3(g), dim3(TB), 0, s, n, k, nb, pos, vel, nbt, soa);
    SphView v{soa, soa + n, soa + 2 * size_t(n), soa + 3 * size_t(n), soa + 4 * size_t(n),
              soa + 5 * size_t(n), mass, u, h, n};
    hipLaunchKernelGGL(k_pack, dim3(g), dim3(TB), 0, s, v, gamma, rec, a4);
    const int gl = gf_div_up((long long)n * SL, TB);
    hipLaunchKernelGGL(k_density, dim3(gl), dim3(TB), 0, s, n, nbt, k, u, gamma, a4, rec, rr);
    hipLaunchKernelGGL(k_force, dim3(gl), dim3(TB), 0, s, n, nbt, k, rec, alpha, beta, nullptr, acc,
                       dudt, dt, divv, rho, nullptr);
    GF_LAUNCH_CHECK();
    return 0;
}

size_t gf_sph_table_scratch_bytes(int n, int k) {
    return sizeof(int) * (size_t(k) * n + 2) + sizeof(double) * (6 + 16 + 1 + 4) * size_t(n) + 2048;
}

int gf_sph_export_knn(GfSph& S, int kq, int64_t* nb, double* dist, hipStream_t s) {
    const int n = S.tree.n;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_export_knn, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, kq, S.nb,
                       S.tree.tx, S.tree.ty, S.tree.tz, S.tree.idx, S.inv, nb, dist);
    GF_LAUNCH_CHECK();
    return 0;
}

__global__ __launch_bounds__(TB) void k_scatter_dk(int n, const double* __restrict__ dk,
                                                   const uint32_t* __restrict__ idx,
                                                   double* __restrict__ out) {
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t < n) out[idx[t]] = dk[t];
}

__global__ __launch_bounds__(TB) void k_export_table(int n, int k, const int* __restrict__ nb,
                                                     const uint32_t* __restrict__ idx,
                                                     int64_t* __restrict__ out) {
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= n) return;
    const size_t row = size_t(idx[t]) * k;
    for (int s = 0; s < k; ++s) out[row + s] = int64_t(idx[nb[size_t(s) * n + t]]);
}

// Number density of the smallest gas-tree cell around each particle that
// holds >= kq particles (the plain first-radius estimate), input order.
__global__ __launch_bounds__(TB) void k_cell_density(int n, int kq, const uint32_t* __restrict__ idx,
                                                     const int* __restrict__ leaf_of,
                                                     const int* __restrict__ parent,
                                                     const int* __restrict__ count,
                                                     const double4* __restrict__ box,
                                                     double* __restrict__ out,
                                                     double* __restrict__ out_count) {
    const int t = blockIdx.x * TB + threadIdx.x;
    if (t >= n) return;
    int node = leaf_of[t];
    while (count[node] < kq && parent[node] >= 0) node = parent[node];
    const double side = 2.0 * box[node].w;
    out[idx[t]] = double(count[node]) / (side * side * side);
    if (out_count) out_count[idx[t]] = double(count[node]);
}

int gf_sph_cell_density(GfSph& S, int kq, double* out, double* out_count, hipStream_t s) {
    const int n = S.tree.n;
    if (n == 0) return 0;
    if (sph_reserve(S, n, 0)) return -1;
    GfTree& T = S.tree;
    hipLaunchKernelGGL(k_leaf_of, dim3(std::min(2048, gf_div_up(T.cap_nodes, TB))), dim3(TB), 0, s,
                       T.counters, T.start, T.count, T.leaf, S.leaf_of);
    hipLaunchKernelGGL(k_cell_density, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, kq, T.idx,
                       S.leaf_of, T.parent, T.count, T.box, out, out_count);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_sph_dk_out(GfSph& S, double* dk, hipStream_t s) {
    const int n = S.tree.n;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_scatter_dk, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, S.dk, S.tree.idx, dk);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_sph_table_out(GfSph& S, int k, int64_t* nb, hipStream_t s) {
    const int n = S.tree.n;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_export_table, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, k, S.nb,
                       S.tree.idx, nb);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_knn_from_candidates(int n, int m, const double* pos, const int64_t* cand, int k,
                           int64_t* nb, double* dist, int* ok, hipStream_t s) {
    if (m > CAND_MAX) {
        gf_set_error("knn_from_candidates: %d candidates per row exceed %d", m, CAND_MAX);
        return -1;
    }
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_knn_cand, dim3(gf_div_up(n, 64)), dim3(64), 0, s, n, m, pos, cand, k, nb,
                       dist, ok);
    GF_LAUNCH_CHECK();
    return 0;
}

int gf_sph_kernel_eval(int n, const double* r, const double* h, double* w, double* g,
                       hipStream_t s) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_kernel_eval, dim3(gf_div_up(n, TB)), dim3(TB), 0, s, n, r, h, w, g);
    GF_LAUNCH_CHECK();
    return 0;
}
