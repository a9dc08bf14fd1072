// Barnes-Hut group walk and list evaluation (cpp/nbody.cpp:323-411), and the
// FP64 direct sum (cpp/nbody.cpp:428-441).
//
// One wave64 walks the tree for one group (the outermost nodes holding <= 64
// particles, nbody.cpp:234-243) and evaluates the resulting interaction list
// for every particle of the group.
//   * The walk is a LIFO of node indices in LDS processed 64 nodes at a time
//     (one node per lane).  The multipole acceptance test is the reference's,
//     |com - b|^2 > (open_r + br)^2 with (b, br) the sphere around the
//     group's massive members (nbody.cpp:350-375), evaluated in FP64 on FP64
//     centres of mass, so both evaluation precisions use one interaction set.
//     Accepted nodes append their monopole to the list; opened leaves append
//     their particles; opened internal nodes push their children.
//   * The list lives in LDS and is evaluated whenever it fills, for the
//     group's active members only, compacted onto the lanes.  Small groups
//     split the list over several lanes per particle ("slices"), so a group
//     of P active particles keeps 64 / P lanes busy per particle.  A group
//     whose P leaves many lanes idle is evaluated as two parts (32, 16 or 8
//     particles, then the rest), each with its own slices.
//   * FP32 evaluation works on coordinates relative to the group centre b,
//     rounded to FP32 once per list entry and group; each flush's FP32 partial
//     sum is added to an FP64 accumulator.
//   * FP64 evaluation uses v_rsq_f64 refined by one Newton step.
// A group whose LDS stack would overflow is queued and walked again by a
// second launch whose stacks live in global memory (bounded by 512 entries
// per tree level: the LIFO keeps entries sorted by depth, and all entries of
// one depth were pushed by one batch of 64 nodes with at most 8 children each).
#include <cstddef>
#include <type_traits>

#include "gf_gravity.h"
#include "gf_prim.h"

namespace {

#ifndef WALK_WAVES
#define WALK_WAVES 4
#endif
// with quadrupoles: a 112-entry particle list and 64 quadrupole entries of
// 36 B, so that a wave holds 8 KB of LDS with its 4 KB stack, and (with the
// flushes unrolled less, below) at most 104 VGPRs: sixteen waves of the walk
// then leave a CU room for one block of the neighbour pass, which the
// overlapped step runs beside it (alone, the walk is 4 % slower than with a
// 192-entry list and 114 VGPRs; in the overlapped step 3 % faster)
#ifndef WALK_LCAP_QUAD
#define WALK_LCAP_QUAD 112
#endif
#ifndef WALK_QCAP
#define WALK_QCAP 64
#endif
#ifndef WALK_SCAP
#define WALK_SCAP 1024
#endif
constexpr int SCAP = WALK_SCAP;                     // LDS stack entries
constexpr int LCAP = 256;                           // LDS list entries
constexpr int LCAP_QUAD = WALK_LCAP_QUAD;           // particle list with quadrupoles
constexpr int QCAP = WALK_QCAP;                     // node list with quadrupoles (36 B each)
// one batch of 64 nodes appends up to 64 entries to a list after one flush
static_assert(LCAP >= 64 && LCAP_QUAD >= 64 && QCAP >= 64, "list capacities below a batch");
constexpr int GSCAP = 512 * (GF_KEY_LEVELS + 2);    // global stack entries per block
// persistent one-wave blocks: one per resident slot of the walk on MI355X
// (256 CUs x 16: four waves per SIMD, 10 KB of LDS each); a grid twice that
// size measured 7 % slower on the 1M disk
#ifndef WALK_GRID
#define WALK_GRID 4096
#endif
constexpr int GST_BLOCKS = 256;

// A tree node packed for the walk (64 bytes).  The records are numbered level
// by level (root first, each level in tree order), so the children of a node
// hold consecutive records and a visit reads one 64-byte record, everything
// the opening test, the quadrupole rule and the descent need, in one round of
// loads for a batch of 64 nodes.
struct GNode {
    double4 com;          // centre of mass, mass (FP64)
    double open_r;        // opening radius (FP64)
    int start, count;     // particle range in tree order
    int kid;              // record of the first non-empty child (internal nodes)
    int nkid;             // non-empty children; -1 for a leaf
    float qlen;           // sqrt(3 max |Q_ij| / m): the quadrupole term's share of
                          // the monopole at distance d is at most about (qlen / d)^2
    float qabs4;          // (2.858 |Q|_F)^(1/4): the quadrupole term's acceleration at
                          // distance d is at most (qabs4 / d)^4 (3.5 |Q|_2 / d^4 with
                          // |Q|_2 <= sqrt(2/3) |Q|_F for a traceless symmetric Q)
};
static_assert(sizeof(GNode) == 64, "GNode layout");
static_assert(offsetof(GNode, open_r) == 32 && offsetof(GNode, start) == 40 &&
                  offsetof(GNode, kid) == 48 && offsetof(GNode, nkid) == 52 &&
                  offsetof(GNode, qlen) == 56 && offsetof(GNode, qabs4) == 60,
              "the walk reads GNode as 16-byte words");
// The traceless quadrupole of a record, FP32, read only for a node accepted
// with its quadrupole: a = (xx, yy, xy, xz), yz.
struct GQuad {
    float4 a;
    float yz, pad[3];
};
static_assert(sizeof(GQuad) == 32, "GQuad layout");
typedef int v4i __attribute__((ext_vector_type(4)));

struct WalkArgs {
    const GNode* gn;           // walk records, level by level (record 0 is the root)
    const GQuad* gq;           // their quadrupoles (QUAD)
    const double4* p4;         // (x, y, z, m) per tree position
    const int* start;
    const int* count;
    const int8_t* leaf;
    const int* child8;
    const double4* com;
    const double* open_r;
    const int* groups;
    const int* counters;
    const double* tx;
    const double* ty;
    const double* tz;
    const double* tm;
    const uint32_t* idx;
    const uint8_t* active;
    double* acc;
    double eps2;
    int* work;                 // [0] LDS-stack launch, [1] global-stack launch
    unsigned long long* inter;
    int* ovf;
    int* ovf_count;            // [0] queued groups, [1] global-stack overflows (error),
                               // [2] record layout errors (error)
    int* gstack;
    int lds_scap;              // LDS stack entries used by the first launch (<= SCAP)
    unsigned long long* inter_nodes;   // the node (monopole or quadrupole) share of inter
    float quad_rtol;           // QUAD: 1 / sqrt(tolerance), +inf for 0: a node whose
                               // quadrupole term is below the tolerance's share of its
                               // monopole at the group's nearest point is a monopole
    float quad_atol;           // QUAD, 0 = off: a node whose quadrupole term at the
                               // group's nearest point is below quad_atol times the
                               // smallest |acc| the group's active members hold on entry
                               // (the previous step's) is a monopole; groups without
                               // that history (a zero or non-finite row, more than 64
                               // members) keep the quad_rtol rule
};

template <bool F64>
struct Acc;

template <>
struct Acc<false> {
    double sx = 0, sy = 0, sz = 0;      // FP64 sum of the flushes' FP32 partial sums
    float px, py, pz;
    // U entries a step: 4 in the monopole walk, 2 in the quadrupole walk,
    // whose register budget is smaller (the entry order of a lane's sum is
    // the same)
    template <int U = 4>
    __device__ void flush(const float4* L, int ln, int sl, int S, float eps2) {
        float ax = 0, ay = 0, az = 0;
        int j = sl;
        for (; j + (U - 1) * S < ln; j += U * S) {
            float4 e[U];
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = L[j + u * S];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float dx = e[u].x - px, dy = e[u].y - py, dz = e[u].z - pz;
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
                const float inv = __builtin_amdgcn_rsqf(r2);
                const float f = e[u].w * (inv * inv * inv);
                ax = fmaf(f, dx, ax);
                ay = fmaf(f, dy, ay);
                az = fmaf(f, dz, az);
            }
        }
        for (; j < ln; j += S) {
            const float4 e = L[j];
            const float dx = e.x - px, dy = e.y - py, dz = e.z - pz;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
            const float inv = __builtin_amdgcn_rsqf(r2);
            const float f = e.w * (inv * inv * inv);
            ax = fmaf(f, dx, ax);
            ay = fmaf(f, dy, ay);
            az = fmaf(f, dz, az);
        }
        sx += double(ax);
        sy += double(ay);
        sz += double(az);
    }
    // Node entries with quadrupoles: e = (x, y, z, m) relative to the group
    // centre, qa = (xx, yy, zz, xy), qb = (xz, yz, -, -).  With d = node -
    // particle and r^2 = |d|^2 + eps^2, the acceleration of the potential
    // -m/r - (1/2) d.Q.d / r^5 is  a = (m/r^3 + (5/2) d.Q.d / r^7) d - Q.d / r^5.
    // qa = (Qxx, Qyy, Qxy, Qxz), qyz = Qyz; Qzz = -(Qxx + Qyy) (traceless)
    __device__ static void quad_term(const float4& e, const float4& qa, float qyz, float px,
                                     float py, float pz, float eps2, float& ax, float& ay,
                                     float& az) {
        const float dx = e.x - px, dy = e.y - py, dz = e.z - pz;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
        const float inv = __builtin_amdgcn_rsqf(r2);
        const float inv2 = inv * inv;
        const float inv3 = inv * inv2;
        const float inv5 = inv3 * inv2;
        const float qzz = -(qa.x + qa.y);
        const float qx = fmaf(qa.x, dx, fmaf(qa.z, dy, qa.w * dz));
        const float qy = fmaf(qa.z, dx, fmaf(qa.y, dy, qyz * dz));
        const float qz = fmaf(qa.w, dx, fmaf(qyz, dy, qzz * dz));
        const float dqd = fmaf(dx, qx, fmaf(dy, qy, dz * qz));
        const float f = fmaf(2.5f * dqd, inv5 * inv2, e.w * inv3);
        ax = fmaf(f, dx, fmaf(-inv5, qx, ax));
        ay = fmaf(f, dy, fmaf(-inv5, qy, ay));
        az = fmaf(f, dz, fmaf(-inv5, qz, az));
    }
    __device__ void flush_quad(const float4* E, const float4* QA, const float* QB, int ln, int sl,
                               int S, float eps2) {
        float ax = 0, ay = 0, az = 0;
        for (int j = sl; j < ln; j += S) quad_term(E[j], QA[j], QB[j], px, py, pz, eps2, ax, ay, az);
        sx += double(ax);
        sy += double(ay);
        sz += double(az);
    }
    __device__ double3 total() const { return make_double3(sx, sy, sz); }
};

template <>
struct Acc<true> {
    double ax = 0, ay = 0, az = 0;
    double px, py, pz;
    __device__ void flush(const double4* L, int ln, int sl, int S, double eps2) {
        for (int j = sl; j < ln; j += S) {
            const double4 e = L[j];
            const double dx = e.x - px, dy = e.y - py, dz = e.z - pz;
            const double r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
            double y = __builtin_amdgcn_rsq(r2);
            y = y * fma(-0.5 * r2, y * y, 1.5);
            const double f = e.w * (y * y * y);
            ax = fma(f, dx, ax);
            ay = fma(f, dy, ay);
            az = fma(f, dz, az);
        }
    }
    __device__ double3 total() const { return make_double3(ax, ay, az); }
};

// A chunk of P <= 64 particles is evaluated in one part or in two.  A part of
// q particles runs 64/q slices over the list, so its pass costs 1/(64/q) of a
// full-wave pass: one part of 40 particles leaves 24 lanes idle for a whole
// pass, where parts of 32 and 8 take 1/2 + 1/8 of one.  Returns the size of
// the first part (32, 16 or 8), or P when a second part would not save at
// least 10 %.
__device__ __forceinline__ int chunk_first_part(int P) {
    int best = P, bnum = 1, bden = 64 / P;          // cost bnum / bden passes
    for (int q = 32; q >= 8; q >>= 1) {
        if (P <= q) continue;
        const int sa = 64 / q, sb = 64 / (P - q);
        if (10 * (sa + sb) * bden < 9 * sa * sb * bnum) {
            best = q;
            bnum = sa + sb;
            bden = sa * sb;
        }
    }
    return best;
}

// Position of set bit number s (counted from 0) of m, for s < popcount(m):
// the largest p with fewer than s + 1 set bits below it.
__device__ __forceinline__ int nth_set_bit(uint64_t m, int s) {
    int pos = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const uint64_t below = m & ((uint64_t(1) << (pos + step)) - 1);   // pos + step <= 63
        if (__popcll(below) <= s) pos += step;
    }
    return pos;
}

// QUAD (FP32 only): accepted nodes carry their quadrupole and go to a list of
// their own; particles keep the monopole list.
template <bool F64, bool GST, bool QUAD>
__global__ __launch_bounds__(64, WALK_WAVES) void k_walk(WalkArgs a) {
    // group sphere and acceptance test: the operations as written (no FMA
    // contraction), so the interaction set is the one the FP64 oracle forms
#pragma clang fp contract(off)
    static_assert(!(F64 && QUAD), "quadrupoles are evaluated in FP32");
    using Ent = typename std::conditional<F64, double4, float4>::type;
    constexpr int LC = QUAD ? LCAP_QUAD : LCAP;
    constexpr int QC = QUAD ? QCAP : 1;
    __shared__ int s_stack[GST ? 1 : SCAP];
    __shared__ Ent s_list[LC];
    __shared__ float4 s_qe[QC], s_qa[QC];
    __shared__ float s_qb[QC];
    // the slice reduction reuses the list memory once the list is evaluated
    static_assert(sizeof(Ent) * LC >= 3 * 64 * sizeof(double), "reduction space");
    double(*s_red)[64] = reinterpret_cast<double(*)[64]>(s_list);
    const int lane = __lane_id();
    int* stack = GST ? a.gstack + size_t(blockIdx.x) * GSCAP : s_stack;
    const int scap = GST ? GSCAP : a.lds_scap;
    const int ngroups = GST ? a.ovf_count[0] : a.counters[1];
    int* cursor = a.work + (GST ? 1 : 0);
    using Real = typename std::conditional<F64, double, float>::type;
    const Real eps2r = Real(a.eps2);     // FP32: float(eps * eps), as nbody.cpp:330

    while (true) {
        int gi = 0;
        if (lane == 0) gi = atomicAdd(cursor, 1);
        gi = __shfl(gi, 0, 64);
        if (gi >= ngroups) break;
        const int G = a.groups[GST ? a.ovf[gi] : gi];
        const int b = a.start[G], np = a.count[G];

        // activity and the sphere around the massive members (FP64)
        bool any_act = false;
        bool no_hist = false;           // QUAD, relative rule: an active member whose
        double amin2 = INFINITY;        // entry |acc|^2 is zero or not finite; the least
        double x0 = 1e300, y0 = 1e300, z0 = 1e300, x1 = -1e300, y1 = -1e300, z1 = -1e300;
        for (int c0 = 0; c0 < np; c0 += 64) {
            const int i = b + c0 + lane;
            const bool in = c0 + lane < np;
            bool act = false;
            if (in) {
                act = a.active ? a.active[a.idx[i]] != 0 : true;
                if constexpr (QUAD) {
                    if (act && a.quad_atol > 0.0f) {
                        const double* o = a.acc + 3 * size_t(a.idx[i]);
                        const double a2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
                        if (a2 > 0.0 && a2 < INFINITY)
                            amin2 = fmin(amin2, a2);
                        else
                            no_hist = true;
                    }
                }
                const double4 q = a.p4[i];
                if (q.w > 0) {
                    const double x = q.x, y = q.y, z = q.z;
                    x0 = fmin(x0, x); x1 = fmax(x1, x);
                    y0 = fmin(y0, y); y1 = fmax(y1, y);
                    z0 = fmin(z0, z); z1 = fmax(z1, z);
                }
            }
            any_act |= __ballot(act) != 0;
        }
        if (!any_act) continue;
        x0 = gf_wave_min_d(x0); y0 = gf_wave_min_d(y0); z0 = gf_wave_min_d(z0);
        x1 = gf_wave_max_d(x1); y1 = gf_wave_max_d(y1); z1 = gf_wave_max_d(z1);
        if (x0 > x1) continue;                                   // massless group
        const double bx = 0.5 * (x0 + x1), by = 0.5 * (y0 + y1), bz = 0.5 * (z0 + z1);
        double br2 = 0;
        for (int c0 = 0; c0 < np; c0 += 64) {
            const int i = b + c0 + lane;
            const double4 q = c0 + lane < np ? a.p4[i] : make_double4(0, 0, 0, 0);
            if (c0 + lane < np && q.w > 0) {
                const double dx = q.x - bx, dy = q.y - by, dz = q.z - bz;
                br2 = fmax(br2, dx * dx + dy * dy + dz * dz);
            }
        }
        const double br = sqrt(gf_wave_max_d(br2));
        // the quadrupole rule of this group: relative to the members' entry
        // accelerations where every active member has one (and the group is
        // one chunk, so no member's row is rewritten before the test), else
        // the share of the node's own monopole
        bool qrel = false;
        float qfac = a.quad_rtol;
        if constexpr (QUAD) {
            if (a.quad_atol > 0.0f && np <= 64) {
                const bool bad = __ballot(no_hist) != 0;
                const double m2 = gf_wave_min_d(amin2);
                if (!bad && m2 < INFINITY) {
                    qrel = true;
                    qfac = float(1.0 / sqrt(sqrt(double(a.quad_atol) * sqrt(m2))));
                }
            }
        }

        // interactions of the whole group, counted once it has finished: a
        // group of several 64-particle chunks whose later chunk overflows the
        // LDS stack is redone from its first chunk by the global-stack walk
        unsigned long long g_len = 0, g_nodes = 0;
        bool g_ovf = false;
        for (int c0 = 0; c0 < np; c0 += 64) {
            const int Pc = min(64, np - c0);
            bool chunk_act = false;
            if (lane < Pc) chunk_act = a.active ? a.active[a.idx[b + c0 + lane]] != 0 : true;
            const uint64_t amask = __ballot(chunk_act);
            const int n_act = __popcll(amask);
            if (n_act == 0) continue;
            // only the active members are evaluated: slot s holds the chunk's
            // s-th active particle (slot s is particle s when all are active);
            // lane s keeps slot s's position in the chunk
            const int P = n_act;
            const int slot_pos = nth_set_bit(amask, min(lane, P - 1));
            // parts A (slots [0, Pa)) and B ([Pa, P), possibly empty): lane l
            // evaluates slice l / Pa of slot l % Pa, and slice l / Pb of slot
            // Pa + l % Pb
            const int Pa = chunk_first_part(P), Pb = P - Pa;
            const int Sa = 64 / Pa, Sb = Pb > 0 ? 64 / Pb : 0;
            const int sla = lane / Pa, slb = Pb > 0 ? lane / Pb : 64;
            const bool evla = sla < Sa, evlb = slb < Sb;
            Acc<F64> A, B;
            {
                const double4 q = a.p4[b + c0 + __shfl(slot_pos, lane % Pa, 64)];
                const double4 r =
                    a.p4[b + c0 + __shfl(slot_pos, Pb > 0 ? Pa + lane % Pb : lane % Pa, 64)];
                if constexpr (F64) {
                    A.px = q.x; A.py = q.y; A.pz = q.z;
                    B.px = r.x; B.py = r.y; B.pz = r.z;
                } else {
                    A.px = float(q.x - bx); A.py = float(q.y - by); A.pz = float(q.z - bz);
                    B.px = float(r.x - bx); B.py = float(r.y - by); B.pz = float(r.z - bz);
                }
            }
            auto eval_list = [&](int len) {
                if constexpr (F64) {
                    if (evla) A.flush(s_list, len, sla, Sa, eps2r);
                    if (evlb) B.flush(s_list, len, slb, Sb, eps2r);
                } else {
                    constexpr int U = QUAD ? 2 : 4;
                    if (evla) A.template flush<U>(s_list, len, sla, Sa, eps2r);
                    if (evlb) B.template flush<U>(s_list, len, slb, Sb, eps2r);
                }
            };
            auto eval_quad = [&](int len) {
                if constexpr (QUAD) {
                    if (evla) A.flush_quad(s_qe, s_qa, s_qb, len, sla, Sa, eps2r);
                    if (evlb) B.flush_quad(s_qe, s_qa, s_qb, len, slb, Sb, eps2r);
                }
            };

            int sn = 1, ln = 0, nq = 0;
            long long total_len = 0, total_nodes = 0;
            bool overflow = false;
            if (lane == 0) stack[0] = 0;
            __syncthreads();
            while (sn > 0) {
                const int take = min(64, sn);
                sn -= take;
                const int node = lane < take ? stack[sn + lane] : -1;
                __syncthreads();
                // every lane loads a record (the root where it has no node),
                // so the visit is free of branches; the empty asm keeps the
                // four loads together, where the compiler would otherwise
                // move the children into the paths that use them, one
                // dependent round trip each
                const bool valid = node >= 0;
                const v4i* r = reinterpret_cast<const v4i*>(a.gn + (valid ? node : 0));
                v4i w0 = r[0], w1 = r[1], w2 = r[2], w3 = r[3];
                asm volatile("" : "+v"(w0), "+v"(w1), "+v"(w2), "+v"(w3));
                const double4 cm = make_double4(__hiloint2double(w0.y, w0.x), __hiloint2double(w0.w, w0.z),
                                                __hiloint2double(w1.y, w1.x), __hiloint2double(w1.w, w1.z));
                const double orr = __hiloint2double(w2.y, w2.x);
                const double dx = cm.x - bx, dy = cm.y - by, dz = cm.z - bz;
                const double t = orr + br;
                const double d2 = dx * dx + dy * dy + dz * dz;
                const bool far = d2 > t * t;
                const bool mono = valid && far;                 // accepted: one list entry
                const bool opened = valid && !far;
                const int first = w2.z;
                const int pc = opened && w3.y < 0 ? w2.w : 0;  // an opened leaf's particles
                const int push = opened && w3.y > 0 ? w3.y : 0;   // an opened node's children
                bool small_q = false;
                if constexpr (QUAD) {
                    // the quadrupole term's share at the group's nearest
                    // point, about (qlen / (d - br))^2, is below the
                    // tolerance iff qlen / sqrt(tolerance) + br < d; with the
                    // relative rule, its bound (qabs4 / (d - br))^4 is below
                    // quad_atol |acc|_min iff qabs4 qfac + br < d (estimates:
                    // FP32 is enough)
                    const float fld = __int_as_float(qrel ? w3.w : w3.z);
                    const float x = fmaf(fld, qfac, float(br));
                    small_q = x * x < float(d2);
                }
                // push children (lane order keeps the stack sorted by depth)
                const int pinc = gf_wave_incl_scan(push);
                const int ptot = __builtin_amdgcn_readlane(pinc, 63);
                if (sn + ptot > scap) {
                    overflow = true;
                    break;
                }
                {
                    // the children's records are consecutive, in tree order
                    const int o = sn + pinc - push;
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                        if (k < push) stack[o + k] = w3.x + k;
                }
                sn += ptot;
                // append this batch's entries, evaluating the list whenever it
                // is full: the accepted nodes' monopoles (written by their
                // lanes), then the opened leaves' particles, one per lane
                const uint64_t mmask = __ballot(mono);
                const int nmono = __popcll(mmask);
                int qslot = -1;                                // QUAD: this lane's tensor slot
                float4 rqa = make_float4(0, 0, 0, 0);          // its (xx, yy, xy, xz), yz
                float rqb = 0.0f;
                if (nmono > 0) {
                    if constexpr (QUAD) {
                        // nodes whose quadrupole term is negligible join the
                        // particle list as monopoles, the others the
                        // quadrupole list
                        const bool quad = mono && !small_q;
                        const uint64_t qmask = __ballot(quad);
                        const uint64_t smask = mmask & ~qmask;
                        const int nquad = __popcll(qmask), nsm = __popcll(smask);
                        total_nodes += nquad;
                        if (nquad > 0) {
                            if (nq + nquad > QC) {
                                __syncthreads();
                                eval_quad(nq);
                                total_len += nq;
                                nq = 0;
                                __syncthreads();
                            }
                            if (quad) {
                                // the tensor is loaded now and lands in its
                                // slot after this batch's particles: no flush
                                // of the quadrupole list comes in between
                                rqa = a.gq[node].a;
                                rqb = a.gq[node].yz;
                                qslot = nq + __popcll(qmask & gf_lanemask_lt());
                                s_qe[qslot] = make_float4(float(cm.x - bx), float(cm.y - by),
                                                          float(cm.z - bz), float(cm.w));
                            }
                            nq += nquad;
                        }
                        if (nsm > 0) {
                            if (ln + nsm > LC) {
                                __syncthreads();
                                eval_list(ln);
                                total_len += ln;
                                ln = 0;
                                __syncthreads();
                            }
                            if (mono && small_q) {
                                const int slot = ln + __popcll(smask & gf_lanemask_lt());
                                s_list[slot] = make_float4(float(cm.x - bx), float(cm.y - by),
                                                           float(cm.z - bz), float(cm.w));
                            }
                            ln += nsm;
                        }
                    } else {
                        total_nodes += nmono;
                        if (ln + nmono > LC) {
                            __syncthreads();
                            eval_list(ln);
                            total_len += ln;
                            ln = 0;
                            __syncthreads();
                        }
                        if (mono) {
                            const int slot = ln + __popcll(mmask & gf_lanemask_lt());
                            if constexpr (F64) {
                                s_list[slot] = cm;
                            } else {
                                s_list[slot] = make_float4(float(cm.x - bx), float(cm.y - by),
                                                           float(cm.z - bz), float(cm.w));
                            }
                        }
                        ln += nmono;
                    }
                }
                const int qinc = gf_wave_incl_scan(pc);
                const int ptot2 = __builtin_amdgcn_readlane(qinc, 63);
                for (int e0 = 0; e0 < ptot2;) {
                    if (ln == LC) {
                        __syncthreads();
                        eval_list(ln);
                        total_len += ln;
                        ln = 0;
                        __syncthreads();
                    }
                    const int m = min(min(64, LC - ln), ptot2 - e0);
                    const int e = e0 + lane;
                    // owner lane of entry e: the number of lanes whose
                    // inclusive count is <= e (all lanes take part)
                    int own = 0;
#pragma unroll
                    for (int step = 32; step > 0; step >>= 1)
                        if (__shfl(qinc, own + step - 1, 64) <= e) own += step;
                    const int ofirst = __shfl(first, own, 64);
                    const int obase = __shfl(qinc - pc, own, 64);
                    if (lane < m) {
                        const double4 q = a.p4[ofirst + (e - obase)];
                        if constexpr (F64) {
                            s_list[ln + lane] = q;
                        } else {
                            s_list[ln + lane] = make_float4(float(q.x - bx), float(q.y - by),
                                                            float(q.z - bz), float(q.w));
                        }
                    }
                    ln += m;
                    e0 += m;
                }
                if constexpr (QUAD) {
                    if (qslot >= 0) {
                        s_qa[qslot] = rqa;
                        s_qb[qslot] = rqb;
                    }
                }
                __syncthreads();
            }
            if (overflow) {
                if (!GST) {
                    if (lane == 0) a.ovf[atomicAdd(&a.ovf_count[0], 1)] = gi;
                } else if (lane == 0) {
                    atomicAdd(&a.ovf_count[1], 1);
                }
                __syncthreads();
                g_ovf = true;
                break;                    // the whole group is redone by the queued walk
            }
            if (ln > 0) {
                __syncthreads();
                eval_list(ln);
                total_len += ln;
                ln = 0;
            }
            if constexpr (QUAD) {
                if (nq > 0) {
                    __syncthreads();
                    eval_quad(nq);
                    total_len += nq;
                    nq = 0;
                }
            }
            // reduce the slices of each particle in FP64, part A then part B
            for (int part = 0; part < (Pb > 0 ? 2 : 1); ++part) {
                const int Pp = part ? Pb : Pa, Sp = part ? Sb : Sa, p0 = part ? Pa : 0;
                double3 v = make_double3(0, 0, 0);
                if (part ? evlb : evla) v = part ? B.total() : A.total();
                __syncthreads();
                s_red[0][lane] = v.x;
                s_red[1][lane] = v.y;
                s_red[2][lane] = v.z;
                const int ppos = __shfl(slot_pos, min(p0 + lane, P - 1), 64);
                __syncthreads();
                if (lane < Pp) {
                    double sx = 0, sy = 0, sz = 0;
                    for (int q = 0; q < Sp; ++q) {
                        sx += s_red[0][q * Pp + lane];
                        sy += s_red[1][q * Pp + lane];
                        sz += s_red[2][q * Pp + lane];
                    }
                    const size_t o = 3 * size_t(a.idx[b + c0 + ppos]);
                    a.acc[o] = sx;
                    a.acc[o + 1] = sy;
                    a.acc[o + 2] = sz;
                }
            }
            __syncthreads();
            g_len += (unsigned long long)total_len * (unsigned long long)n_act;
            g_nodes += (unsigned long long)total_nodes * (unsigned long long)n_act;
        }
        if (!g_ovf && lane == 0) {
            atomicAdd(a.inter, g_len);
            atomicAdd(a.inter_nodes, g_nodes);
        }
    }
}

// Keys of the records' order: the node depth (255 past the node count), with
// the node ids as values; one stable pass on the low byte sorts them level by
// level, each level in tree order.
__global__ __launch_bounds__(256) void k_rec_keys(const int* __restrict__ counters,
                                                 const int8_t* __restrict__ depth, int cap,
                                                 uint64_t* __restrict__ key,
                                                 uint32_t* __restrict__ val) {
    const int nn = counters[0];
    for (int id = blockIdx.x * 256 + threadIdx.x; id < cap; id += gridDim.x * 256) {
        key[id] = id < nn ? uint64_t(uint8_t(depth[id])) : 255u;
        val[id] = uint32_t(id);
    }
}

__global__ __launch_bounds__(256) void k_rec_of(const int* __restrict__ counters,
                                               const uint32_t* __restrict__ order,
                                               int* __restrict__ rec_of) {
    const int nn = counters[0];
    for (int r = blockIdx.x * 256 + threadIdx.x; r < nn; r += gridDim.x * 256)
        rec_of[order[r]] = r;
}

// Record r holds node order[r].  In tree order (depth first, children in
// octant order) the non-empty children of a node are consecutive among the
// nodes of their depth, so they hold consecutive records; a node for which
// that fails is counted in *bad (an error the caller reports).
__global__ __launch_bounds__(256) void k_gnode_pack(const int* __restrict__ counters,
                                                   const uint32_t* __restrict__ order,
                                                   const int* __restrict__ rec_of,
                                                   const double4* __restrict__ com,
                                                   const double* __restrict__ open_r,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ count,
                                                   const int8_t* __restrict__ leaf,
                                                   const int* __restrict__ child8,
                                                   const double* __restrict__ quad,
                                                   GNode* __restrict__ gn,
                                                   GQuad* __restrict__ gq, int* bad) {
    const int nn = counters[0];
    for (int r = blockIdx.x * 256 + threadIdx.x; r < nn; r += gridDim.x * 256) {
        const int id = int(order[r]);
        GNode g;
        g.com = com[id];
        g.open_r = open_r[id];
        g.start = start[id];
        g.count = count[id];
        g.kid = 0;
        g.nkid = -1;
        if (!leaf[id]) {
            int nc = 0, k0 = -1;
            bool ok = true;
            for (int k = 0; k < 8; ++k) {
                const int c = child8[8 * size_t(id) + k];
                if (c < 0) continue;
                const int rc = rec_of[c];
                if (nc == 0) k0 = rc;
                ok &= rc == k0 + nc;
                ++nc;
            }
            if (!ok) atomicAdd(bad, 1);
            g.kid = nc > 0 ? k0 : 0;
            g.nkid = nc;
        }
        double qmax = 0.0, qf2 = 0.0;
        GQuad q4;
        q4.a = make_float4(0, 0, 0, 0);
        q4.yz = q4.pad[0] = q4.pad[1] = q4.pad[2] = 0.0f;
        if (quad) {
            const double* Q = quad + 6 * size_t(id);   // (xx, yy, zz, xy, xz, yz)
            for (int k = 0; k < 6; ++k) qmax = fmax(qmax, fabs(Q[k]));
            qf2 = Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2] +
                  2.0 * (Q[3] * Q[3] + Q[4] * Q[4] + Q[5] * Q[5]);
            q4.a = make_float4(float(Q[0]), float(Q[1]), float(Q[3]), float(Q[4]));
            q4.yz = float(Q[5]);
            gq[r] = q4;
        }
        g.qlen = g.com.w > 0 ? float(sqrt(3.0 * qmax / g.com.w)) : 0.0f;
        g.qabs4 = float(sqrt(sqrt(2.858 * sqrt(qf2))));
        gn[r] = g;
    }
}

__global__ __launch_bounds__(256) void k_gp4_pack(int n, const double* __restrict__ tx,
                                                 const double* __restrict__ ty,
                                                 const double* __restrict__ tz,
                                                 const double* __restrict__ tm,
                                                 double4* __restrict__ p4) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) p4[t] = make_double4(tx[t], ty[t], tz[t], tm[t]);
}

// The walk's cursors and overflow counts; with records, also the record
// layout error count that the packing of the records adds to.
__global__ void k_walk_reset(int* work, int* ovf_count, int records) {
    if (threadIdx.x == 0) {
        work[0] = 0;
        work[1] = 0;
        ovf_count[0] = 0;
        ovf_count[1] = 0;
        if (records) ovf_count[2] = 0;
    }
}

// ------------------------------------------------------------ direct sum
constexpr int DT = 256;

__global__ __launch_bounds__(DT) void k_direct(int n, const double* __restrict__ pos,
                                               const double* __restrict__ mass, int nt,
                                               const int64_t* __restrict__ tgt, double eps2,
                                               int per_split, double* __restrict__ partial) {
    __shared__ double4 src[DT];
    const int t = blockIdx.x * DT + threadIdx.x;
    const int split = blockIdx.y;
    const bool ok = t < nt;
    const long long i = ok ? (tgt ? tgt[t] : t) : 0;
    const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    double ax = 0, ay = 0, az = 0;
    const int j0 = split * per_split, j1 = min(n, j0 + per_split);
    for (int jb = j0; jb < j1; jb += DT) {
        const int j = jb + threadIdx.x;
        __syncthreads();
        src[threadIdx.x] = j < j1 ? make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], mass[j])
                                  : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int m = min(DT, j1 - jb);
        for (int k = 0; k < m; ++k) {
            const double4 s = src[k];
            const double dx = s.x - x, dy = s.y - y, dz = s.z - z;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            const double f = s.w / (r2 * sqrt(r2));
            ax += f * dx;
            ay += f * dy;
            az += f * dz;
        }
    }
    if (ok) {
        const size_t o = 3 * (size_t(split) * nt + t);
        partial[o] = ax;
        partial[o + 1] = ay;
        partial[o + 2] = az;
    }
}

__global__ __launch_bounds__(DT) void k_direct_reduce(int nt, int nsplit,
                                                      const double* __restrict__ partial,
                                                      double* __restrict__ acc) {
    const int t = blockIdx.x * DT + threadIdx.x;
    if (t >= nt) return;
    double ax = 0, ay = 0, az = 0;
    for (int s = 0; s < nsplit; ++s) {
        const size_t o = 3 * (size_t(s) * nt + t);
        ax += partial[o];
        ay += partial[o + 1];
        az += partial[o + 2];
    }
    acc[3 * size_t(t)] = ax;
    acc[3 * size_t(t) + 1] = ay;
    acc[3 * size_t(t) + 2] = az;
}

// Softened potential phi_i = -sum_{j != i} m_j / sqrt(r_ij^2 + eps^2), FP64,
// for the targets t0 .. t0 + nt - 1 of the n particles (the energy of a run;
// no counterpart in the reference).  partial[split][t - t0].
__global__ __launch_bounds__(DT) void k_potential(int n, int t0, int nt,
                                                  const double* __restrict__ pos,
                                                  const double* __restrict__ mass, double eps2,
                                                  int per_split, double* __restrict__ partial) {
    __shared__ double4 src[DT];
    const int tt = blockIdx.x * DT + threadIdx.x;
    const int t = t0 + tt;
    const int split = blockIdx.y;
    const bool ok = tt < nt;
    const int ti = ok ? t : 0;
    const double x = pos[3 * size_t(ti)], y = pos[3 * size_t(ti) + 1], z = pos[3 * size_t(ti) + 2];
    double phi = 0;
    const int j0 = split * per_split, j1 = min(n, j0 + per_split);
    for (int jb = j0; jb < j1; jb += DT) {
        const int j = jb + threadIdx.x;
        __syncthreads();
        src[threadIdx.x] = j < j1 ? make_double4(pos[3 * size_t(j)], pos[3 * size_t(j) + 1],
                                                 pos[3 * size_t(j) + 2], mass[j])
                                  : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int m = min(DT, j1 - jb);
        for (int k = 0; k < m; ++k) {
            const double4 q = src[k];
            const double dx = q.x - x, dy = q.y - y, dz = q.z - z;
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            double inv = __builtin_amdgcn_rsq(r2);
            inv = inv * fma(-0.5 * r2, inv * inv, 1.5);
            phi -= (jb + k != t) ? q.w * inv : 0.0;
        }
    }
    if (ok) partial[size_t(split) * nt + tt] = phi;
}

__global__ __launch_bounds__(DT) void k_potential_reduce(int t0, int nt, int nsplit,
                                                         const double* __restrict__ partial,
                                                         double* __restrict__ phi) {
    const int tt = blockIdx.x * DT + threadIdx.x;
    if (tt >= nt) return;
    double s = 0;
    for (int k = 0; k < nsplit; ++k) s += partial[size_t(k) * nt + tt];
    phi[t0 + tt] = s;
}

// Targets per potential launch: 2^18 targets of 4M sources is ~1e12 pair
// terms, well under a second, so no launch holds the device for long.
constexpr int POT_CHUNK = 1 << 18;

int direct_splits(int n, int nt) {
    const int tb = gf_div_up(nt, DT);
    const int want = std::max(1, 2048 / std::max(1, tb));
    return std::max(1, std::min(want, gf_div_up(n, DT)));
}

}  // namespace

int gf_walk_reserve(GfWalkScratch& W, int max_groups) {
    if (!W.work) {
        GF_HIP(hipMalloc(&W.work, 16 * sizeof(int)));
        GF_HIP(hipMalloc(&W.inter, 2 * sizeof(unsigned long long)));
        W.inter_nodes = W.inter + 1;
        GF_HIP(hipMalloc(&W.ovf_count, 4 * sizeof(int)));
        GF_HIP(hipMalloc(&W.gstack, sizeof(int) * size_t(GST_BLOCKS) * GSCAP));
        GF_HIP(hipMemset(W.inter, 0, 2 * sizeof(unsigned long long)));
    }
    if (max_groups > W.cap_groups) {
        if (W.ovf) GF_HIP(hipFree(W.ovf));
        GF_HIP(hipMalloc(&W.ovf, sizeof(int) * size_t(std::max(1, max_groups))));
        W.cap_groups = max_groups;
    }
    return 0;
}

// Packed copies of the tree for the walk (sized for cap_nodes nodes and cap_n particles).
static int walk_pack_reserve(GfWalkScratch& W, int cap_nodes, int cap_n) {
    if (cap_nodes > W.cap_gnode) {
        void** bufs[] = {&W.gnode, &W.gquad, &W.rkey, &W.rval, &W.rec_of, &W.rsort_tmp};
        for (void** b : bufs) {
            if (*b) GF_HIP(hipFree(*b));
            *b = nullptr;
        }
        const size_t c = size_t(std::max(1, cap_nodes));
        GF_HIP(hipMalloc(&W.gnode, sizeof(GNode) * c));
        GF_HIP(hipMalloc(&W.gquad, sizeof(GQuad) * c));
        GF_HIP(hipMalloc(&W.rkey, 2 * sizeof(uint64_t) * c));
        GF_HIP(hipMalloc(&W.rval, 2 * sizeof(uint32_t) * c));
        GF_HIP(hipMalloc(&W.rec_of, sizeof(int) * c));
        GF_HIP(hipMalloc(&W.rsort_tmp, sizeof(int) * gf_sort_temp_ints(int(c))));
        W.cap_gnode = cap_nodes;
    }
    if (cap_n > W.cap_p4) {
        if (W.p4) GF_HIP(hipFree(W.p4));
        GF_HIP(hipMalloc(&W.p4, sizeof(double4) * size_t(std::max(1, cap_n))));
        W.cap_p4 = cap_n;
    }
    return 0;
}

void gf_walk_free(GfWalkScratch& W) {
    if (W.work) (void)hipFree(W.work);
    if (W.inter) (void)hipFree(W.inter);
    if (W.ovf_count) (void)hipFree(W.ovf_count);
    if (W.gstack) (void)hipFree(W.gstack);
    if (W.ovf) (void)hipFree(W.ovf);
    void* bufs[] = {W.gnode, W.gquad, W.rkey, W.rval, W.rec_of, W.rsort_tmp, W.p4};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    W = GfWalkScratch();
}

int gf_walk_default_blocks() { return WALK_GRID; }

int gf_walk_prepare(const GfTree& T, GfWalkScratch& W, int precision, hipStream_t s) {
    if (T.n == 0) return 0;
    if (!T.built) {
        gf_set_error("gravity walk without a built tree");
        return -1;
    }
    const int quad = precision == GF_FP32_QUAD ? 1 : 0;
    if (W.prep_tree == &T && W.prep_version == T.version && W.prep_quad == quad) return 0;
    if (gf_walk_reserve(W, T.cap_nodes) || walk_pack_reserve(W, T.cap_nodes, T.n)) return -1;
    T.want_quad = quad;
    if (quad && gf_tree_quadrupoles(T, s)) return -1;
    hipLaunchKernelGGL(k_walk_reset, dim3(1), dim3(64), 0, s, W.work, W.ovf_count, 1);
    // the records' order: nodes by depth, each depth in tree order
    const int cn = T.cap_nodes, gb = std::min(2048, gf_div_up(cn, 256));
    uint64_t* rkey = static_cast<uint64_t*>(W.rkey);
    uint32_t* rval = static_cast<uint32_t*>(W.rval);
    hipLaunchKernelGGL(k_rec_keys, dim3(gb), dim3(256), 0, s, T.counters, T.depth, cn, rkey, rval);
    int alt = 0;
    if (gf_radix_sort_pairs(rkey, rval, rkey + cn, rval + cn, cn, 0, 8,
                            static_cast<int*>(W.rsort_tmp), s, &alt))
        return -1;
    const uint32_t* order = alt ? rval + cn : rval;
    hipLaunchKernelGGL(k_rec_of, dim3(gb), dim3(256), 0, s, T.counters, order,
                       static_cast<int*>(W.rec_of));
    hipLaunchKernelGGL(k_gnode_pack, dim3(gb), dim3(256), 0, s, T.counters, order,
                       static_cast<const int*>(W.rec_of), T.com, T.open_r, T.start, T.count,
                       T.leaf, T.child8, precision == GF_FP32_QUAD ? T.quad : nullptr,
                       static_cast<GNode*>(W.gnode), static_cast<GQuad*>(W.gquad),
                       W.ovf_count + 2);
    hipLaunchKernelGGL(k_gp4_pack, dim3(gf_div_up(T.n, 256)), dim3(256), 0, s, T.n, T.tx, T.ty,
                       T.tz, T.tm, reinterpret_cast<double4*>(W.p4));
    GF_LAUNCH_CHECK();
    W.prep_tree = &T;
    W.prep_version = T.version;
    W.prep_quad = quad;
    return 0;
}

int gf_walk(const GfTree& T, GfWalkScratch& W, const uint8_t* active, double* acc, double eps,
            int precision, hipStream_t s) {
    if (T.n == 0) return 0;
    if (gf_walk_prepare(T, W, precision, s)) return -1;
    hipLaunchKernelGGL(k_walk_reset, dim3(1), dim3(64), 0, s, W.work, W.ovf_count, 0);
    WalkArgs a;
    a.gn = static_cast<const GNode*>(W.gnode);
    a.gq = static_cast<const GQuad*>(W.gquad);
    a.p4 = reinterpret_cast<const double4*>(W.p4);
    a.start = T.start; a.count = T.count; a.leaf = T.leaf; a.child8 = T.child8;
    a.com = T.com; a.open_r = T.open_r; a.groups = T.groups; a.counters = T.counters;
    a.tx = T.tx; a.ty = T.ty; a.tz = T.tz; a.tm = T.tm; a.idx = T.idx;
    a.active = active; a.acc = acc; a.eps2 = eps * eps;
    a.work = W.work; a.inter = W.inter; a.ovf = W.ovf; a.ovf_count = W.ovf_count;
    a.gstack = W.gstack;
    a.inter_nodes = W.inter_nodes;
    a.lds_scap = gf_test_limits().walk_stack > 0 ? std::min(SCAP, gf_test_limits().walk_stack) : SCAP;
    a.quad_rtol = W.quad_tol > 0 ? float(1.0 / std::sqrt(double(W.quad_tol))) : INFINITY;
    a.quad_atol = W.quad_rel_tol;
    // the groups are drawn from a queue, so any grid walks them all
    const int WALK_BLOCKS = W.walk_blocks > 0 ? W.walk_blocks : WALK_GRID;
    if (precision == GF_FP64) {
        hipLaunchKernelGGL((k_walk<true, false, false>), dim3(WALK_BLOCKS), dim3(64), 0, s, a);
        hipLaunchKernelGGL((k_walk<true, true, false>), dim3(GST_BLOCKS), dim3(64), 0, s, a);
    } else if (precision == GF_FP32_QUAD) {
        hipLaunchKernelGGL((k_walk<false, false, true>), dim3(WALK_BLOCKS), dim3(64), 0, s, a);
        hipLaunchKernelGGL((k_walk<false, true, true>), dim3(GST_BLOCKS), dim3(64), 0, s, a);
    } else {
        hipLaunchKernelGGL((k_walk<false, false, false>), dim3(WALK_BLOCKS), dim3(64), 0, s, a);
        hipLaunchKernelGGL((k_walk<false, true, false>), dim3(GST_BLOCKS), dim3(64), 0, s, a);
    }
    GF_LAUNCH_CHECK();
    return 0;
}

size_t gf_direct_partial_doubles(int n, int nt) {
    return 3 * size_t(std::max(nt, 1)) * size_t(direct_splits(n, nt));
}

int gf_potential(int n, const double* pos, const double* mass, double* phi, double eps,
                 double* partial, hipStream_t s) {
    if (n <= 0) return 0;
    for (int t0 = 0; t0 < n; t0 += POT_CHUNK) {
        const int nt = std::min(POT_CHUNK, n - t0);
        const int ns = direct_splits(n, nt);
        const int per = gf_div_up(gf_div_up(n, ns), DT) * DT;
        const int nsplit = gf_div_up(n, per);
        hipLaunchKernelGGL(k_potential, dim3(gf_div_up(nt, DT), nsplit), dim3(DT), 0, s, n, t0, nt,
                           pos, mass, eps * eps, per, partial);
        hipLaunchKernelGGL(k_potential_reduce, dim3(gf_div_up(nt, DT)), dim3(DT), 0, s, t0, nt,
                           nsplit, partial, phi);
        GF_LAUNCH_CHECK();
    }
    return 0;
}

size_t gf_potential_partial_doubles(int n) {
    // the largest launch: a full chunk, or the last (shorter, more splits)
    n = std::max(n, 1);
    const int full = std::min(POT_CHUNK, n), last = n - (n - 1) / POT_CHUNK * POT_CHUNK;
    return std::max(size_t(full) * size_t(direct_splits(n, full)),
                    size_t(last) * size_t(direct_splits(n, last)));
}

int gf_direct(int n, const double* pos, const double* mass, int nt, const int64_t* tgt,
              double* acc, double eps, double* partial, hipStream_t s) {
    if (nt <= 0 || n <= 0) return 0;
    const int ns = direct_splits(n, nt);
    const int per = gf_div_up(gf_div_up(n, ns), DT) * DT;
    const int nsplit = gf_div_up(n, per);
    hipLaunchKernelGGL(k_direct, dim3(gf_div_up(nt, DT), nsplit), dim3(DT), 0, s, n, pos, mass,
                       nt, tgt, eps * eps, per, partial);
    hipLaunchKernelGGL(k_direct_reduce, dim3(gf_div_up(nt, DT)), dim3(DT), 0, s, nt, nsplit,
                       partial, acc);
    GF_LAUNCH_CHECK();
    return 0;
}
