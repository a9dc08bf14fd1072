// Barnes-Hut group walk + list evaluation, and the FP64 direct sum (gf_gravity.hip).
#pragma once
#include "gf_tree.h"

// FP32 monopoles, FP64 monopoles, FP32 monopoles + node quadrupoles
enum GfPrecision { GF_FP32 = 0, GF_FP64 = 1, GF_FP32_QUAD = 2 };

struct GfWalkScratch {
    int* work = nullptr;                   // 1 + 8 ints: group cursor(s)
    unsigned long long* inter = nullptr;   // interaction counter ([0] all, [1] node entries)
    unsigned long long* inter_nodes = nullptr;
    int* ovf = nullptr;                    // groups whose LDS stack overflowed
    int* ovf_count = nullptr;
    int* gstack = nullptr;                 // global stacks for those groups
    int cap_groups = 0;
    void* gnode = nullptr;                 // walk records (64 bytes each), level by level
    void* gquad = nullptr;                 // their quadrupoles (32 bytes each)
    void* rkey = nullptr;                  // the records' order: sort keys and node ids,
    void* rval = nullptr;                  // two buffers each
    void* rec_of = nullptr;                // record of each node
    void* rsort_tmp = nullptr;
    void* p4 = nullptr;                    // packed particles (x, y, z, m), tree order
    int cap_gnode = 0, cap_p4 = 0;
    // precision 2: an accepted node whose quadrupole term is estimated below
    // this share of its monopole at the group's nearest point is evaluated
    // as a monopole (0 = every accepted node with its quadrupole)
    float quad_tol = 0.01f;
    // precision 2, 0 = off: an accepted node whose quadrupole term is bounded
    // below this fraction of the smallest |acc| the group's active members
    // hold on entry (the caller's previous accelerations) is evaluated as a
    // monopole; groups without that history use quad_tol
    float quad_rel_tol = 0.0f;
    // persistent one-wave blocks of the walk (0 = the default, gf_walk_default_blocks)
    int walk_blocks = 0;
    // the packed records were made from this tree version, with (1) or
    // without (0) quadrupole estimates (-1: none made)
    int prep_version = -1, prep_quad = -1;
    const GfTree* prep_tree = nullptr;
};

int gf_walk_reserve(GfWalkScratch& W, int max_groups);
// The walk's persistent grid when walk_blocks is 0.
int gf_walk_default_blocks();
void gf_walk_free(GfWalkScratch& W);

// Packs the walk's records of the tree T for a walk at `precision` (node
// records level by level, particles (x, y, z, m), and for precision 2 the
// quadrupoles and their estimates), unless the records of this tree
// version and kind are already held.  gf_walk calls it; a caller may call
// it earlier on the walk's stream, beside other work.
int gf_walk_prepare(const GfTree& T, GfWalkScratch& W, int precision, hipStream_t s);

// Accelerations of the active particles (active: input order, may be null =
// all) from the tree T, written to acc (n, 3) in input order; inactive rows
// and rows of massless groups are left untouched (cpp/nbody.cpp:344-365,
// :404-407).  Adds the interaction count to W.inter.
int gf_walk(const GfTree& T, GfWalkScratch& W, const uint8_t* active, double* acc, double eps,
            int precision, hipStream_t s);

// FP64 direct sum (cpp/nbody.cpp:428-441) for nt targets (tgt: indices, or
// null for all n with nt == n).  partial: 3 * nt * nsplit doubles of scratch
// (gf_direct_partial_doubles).
size_t gf_direct_partial_doubles(int n, int nt);
int gf_direct(int n, const double* pos, const double* mass, int nt, const int64_t* tgt,
              double* acc, double eps, double* partial, hipStream_t s);

// Softened FP64 potential of all n particles by direct summation (the
// energy of a run): phi_i = -sum_{j != i} m_j / sqrt(r_ij^2 + eps^2).
// partial: gf_potential_partial_doubles(n) doubles of scratch.
size_t gf_potential_partial_doubles(int n);
int gf_potential(int n, const double* pos, const double* mass, double* phi, double eps,
                 double* partial, hipStream_t s);
