// Device octree of the Barnes-Hut gravity (gf_tree.hip).
#pragma once
#include "gf_common.h"

// All pointers are device memory.  Particle arrays are indexed by tree
// position (Morton order) unless noted; node arrays by DFS index, the layout
// of cpp/nbody.cpp:176-200 (preorder, children in octant order 0..7).
struct GfTree {
    int n = 0;          // particles the held tree was built for
    int cap_n = 0;      // particle capacity of the buffers
    int cap_nodes = 0;  // node capacity of the buffers
    int built = 0;      // a tree is held
    int version = 0;    // counts builds and refits (the walk's packed records follow it)
    int refit_ok = 0;   // the held tree was built with centres of mass (refit allowed)
    float theta = 0.7f; // opening angle of the held tree (cast to float, nbody.cpp:327)

    // per particle
    uint64_t *key = nullptr, *key_alt = nullptr;
    uint32_t *idx = nullptr, *idx_alt = nullptr;  // tree position -> input index
    int8_t *lcp = nullptr, *d0 = nullptr;
    int *chain = nullptr;                          // n + 1, then node base per position
    double *tx = nullptr, *ty = nullptr, *tz = nullptr, *tm = nullptr;
    double *tdisp = nullptr;                       // displacement since the build
    double *xb = nullptr;                          // positions at the build (input order, AoS)

    // per node
    int *start = nullptr, *count = nullptr, *next = nullptr, *parent = nullptr;
    int8_t *leaf = nullptr, *depth = nullptr;
    int *child8 = nullptr;                         // 8 slots per node, octant order, -1 = empty
    int *dlist = nullptr;                          // nodes above the bottom-up subtrees
    int *dmeta = nullptr;                          // [0] their count (0 between passes)
    double4 *com = nullptr;                        // centre of mass and mass (FP64)
    double *open_r = nullptr;                      // opening radius (FP64)
    double4 *box = nullptr;                        // geometric centre and half-width (FP64)
    double *ndisp = nullptr;                       // largest displacement below the node
    double *quad = nullptr;                        // traceless quadrupole about the centre of
                                                   // mass (xx, yy, zz, xy, xz, yz), FP64
    mutable int quad_ok = 0;                       // quad matches the held centres of mass
    mutable int want_quad = 0;                     // the last walk used quadrupoles: the next
                                                   // centre-of-mass pass computes them too
    int *groups = nullptr;                         // group nodes, DFS order

    // scalars on the device: [0] nodes, [1] groups, [2] cells cut at the key
    // depth with more than GF_LEAF_SIZE particles, [3] spare
    int *counters = nullptr;
    double *root = nullptr;                        // root centre (3) and half-width
    double *bbox_part = nullptr;
    int *scan_tmp = nullptr, *sort_tmp = nullptr;
    // keep_sorted: every build also keeps its Morton-sorted (key, input
    // index), before particles within leaves are put in input order, for
    // subset trees built from it (gf_tree_build_subset)
    int keep_sorted = 0;
    uint64_t* skey = nullptr;
    uint32_t* sidx = nullptr;
    // a subset tree's scratch: the compaction's look-back state over the
    // source tree's n
    int* sub_tmp = nullptr;
    int cap_sub = 0;
};

// Node capacity that always suffices for n particles: at most n leaves and,
// per depth, at most n / (GF_LEAF_SIZE + 1) internal nodes.
int gf_tree_node_bound(int n);

int gf_tree_reserve(GfTree& T, int n);
void gf_tree_free(GfTree& T);

// Builds the tree of pos (n, 3) / mass (n) (device, input order).
// with_com = false skips the centres of mass (a tree used for neighbour search).
// sorted_event (may be null) is recorded on s once the sorted keys are kept.
int gf_tree_build(GfTree& T, int n, const double* pos, const double* mass, float theta,
                  hipStream_t s, bool with_com = true, hipEvent_t sorted_event = nullptr);
// Builds S as the tree of the first m particles of the arrays the tree T
// was last built from (T.keep_sorted set): T's sorted keys restricted to
// them (a stable compaction, no sort of their own; S's root box is T's),
// then S's nodes, leaf order, gathered positions and groups, without
// centres of mass.  pos, mass: the arrays T was built from.  T's build
// must be complete on the device before S's kernels run (the caller orders
// the streams).
int gf_tree_build_subset(GfTree& S, const GfTree& T, int m, const double* pos, const double* mass,
                         hipStream_t s);
// Work items of runs of sibling groups (children of one node, consecutive
// in octant order) of at most cap <= 64 particles in all, every other group
// an item of its own: items[i] = (first group position, groups, first tree
// position, particles), appended in any order after *nitems (device; the
// caller zeroes it first) by atomic increments.
int gf_tree_merge_groups(const GfTree& T, int cap, int* nitems, int4* items, hipStream_t s);
// Quadrupole moments of the held tree (computed once per build or refit, when
// a walk asks for them).
int gf_tree_quadrupoles(const GfTree& T, hipStream_t s);

// Refit (cpp/nbody.cpp:262-305): same cells, new masses, centres of mass and
// opening radii enlarged by the largest displacement since the build.
int gf_tree_refit(GfTree& T, const double* pos, const double* mass, hipStream_t s);
